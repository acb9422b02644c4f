"""Learned Stein control variate for the NVT stress-fluctuation estimator.

New work (no counterpart in the reference). For the canonical distribution
p(x) ~ exp(-beta U(x)) and any smooth vector field g that is periodic in the
box, E_p[div g - beta g . grad U] = 0 (integrate div(g p) over the torus). So
with F = -grad U,

    s_c(x) = sum_i [ div_i g^c_i(x) + beta g^c_i(x) . F_i ],     E[s_c] = 0,

and the controlled estimate mean(y_c - s_c) of a head c has the same
expectation as the plain mean(y_c), for any g. Here g^c_i is a per-atom
3-vector from a small network on radial-basis descriptors of atom i's
neighbourhood (32 Gaussians per neighbour species, 2 hidden tanh layers of
width 64, heads for C11, C12, C44). The network is trained to minimise
Var(y_c - s_c) on frames of the plain run; the plain estimator stays the
judge, and a head whose variance the variate does not cut is switched off.

The device work is hand-written HIP (include/mdelast_cv.h): descriptors and
their derivatives along each atom's own coordinates, and a dense layer that
carries those three tangents forward (Y = act(XW + b), Y'_a = act'(XW + b) T_a W),
with its backward, on bf16 MFMA with fp32 accumulation. PyTorch holds the
device memory and runs the optimiser; ``TangentDense`` is the
torch.autograd.Function over the kernels.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from . import units as U

NB = 32            # radial functions per neighbour species
NF = 2 * NB        # descriptor length (two species)
HIDDEN = 64
NOUT = 16          # 3 heads x 3 components, padded to the MFMA tile
HEADS = ("C11", "C12", "C44")


def _torch():
    import torch
    return torch


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _check(st):
    if st != 0:
        raise nat.MdlError(st, nat.lib().mdl_last_error().decode(errors="replace"))


@dataclass
class Basis:
    rc: float = 6.8          # A, cut-off of the descriptors (the ZnCu EAM cut-off)
    mu0: float = 2.0         # A, centre of the first Gaussian
    dmu: float = 0.15        # A, spacing (32 functions up to 6.65 A)
    width: float = 0.15      # A


def descriptors(pos_A, types, box_A, basis: Basis = Basis()):
    """D [N, 64] and J [3, N, 64] on the device for FP64 positions in A."""
    torch = _torch()
    pos = pos_A.contiguous()
    N = pos.shape[0]
    D = torch.empty((N, NF), device=pos.device, dtype=torch.float32)
    J = torch.empty((3, N, NF), device=pos.device, dtype=torch.float32)
    box = (C.c_double * 3)(*[float(b) for b in box_A])
    _check(nat.lib().mdl_cv_descriptors(_p(pos), _p(types.contiguous()), N, box,
                                        basis.rc, basis.mu0, basis.dmu, basis.width,
                                        _p(D), _p(J), _stream()))
    return D, J


def _gemm(A, W, bias=None, transB=False):
    torch = _torch()
    A = A.contiguous()
    R, K = A.shape
    N = W.shape[0] if transB else W.shape[1]
    Cm = torch.empty((R, N), device=A.device, dtype=torch.float32)
    _check(nat.lib().mdl_cv_gemm(_p(A), _p(W.contiguous()), 1 if transB else 0,
                                 _p(bias.contiguous()) if bias is not None else None,
                                 _p(Cm), R, K, N, _stream()))
    return Cm


def _gemm_tn(A, B, rows=512):
    torch = _torch()
    A, B = A.contiguous(), B.contiguous()
    R, K = A.shape
    N = B.shape[1]
    nchunk = (R + rows - 1) // rows
    work = torch.empty((nchunk, K, N), device=A.device, dtype=torch.float32)
    out = torch.empty((K, N), device=A.device, dtype=torch.float32)
    _check(nat.lib().mdl_cv_gemm_tn(_p(A), _p(B), _p(out), _p(work), R, K, N, rows, _stream()))
    return out


class TangentDense:
    """Y, Yt = act(X W + b), act'(X W + b) (T W): X [M, K], T [3, M, K]."""

    @staticmethod
    def _fn():
        torch = _torch()

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, X, T, W, b, act):
                M, K = X.shape
                N = W.shape[1]
                Z = _gemm(X, W, b)
                Zt = _gemm(T.reshape(3 * M, K), W).reshape(3, M, N)
                Y = torch.empty_like(Z)
                Yt = torch.empty_like(Zt)
                _check(nat.lib().mdl_cv_act_forward(_p(Z), _p(Zt), _p(Y), _p(Yt), M * N, act,
                                                    _stream()))
                ctx.save_for_backward(X, T, W, Z, Zt)
                ctx.act = act
                return Y, Yt

            @staticmethod
            def backward(ctx, gY, gYt):
                X, T, W, Z, Zt = ctx.saved_tensors
                M, K = X.shape
                N = W.shape[1]
                gY = gY.contiguous() if gY is not None else torch.zeros_like(Z)
                gYt = gYt.contiguous() if gYt is not None else torch.zeros_like(Zt)
                gZ = torch.empty_like(Z)
                gZt = torch.empty_like(Zt)
                _check(nat.lib().mdl_cv_act_backward(_p(Z), _p(Zt), _p(gY), _p(gYt), _p(gZ),
                                                     _p(gZt), M * N, ctx.act, _stream()))
                gX = _gemm(gZ, W, transB=True)
                gT = _gemm(gZt.reshape(3 * M, N), W, transB=True).reshape(3, M, K)
                gW = _gemm_tn(torch.cat([X, T.reshape(3 * M, K)]),
                              torch.cat([gZ, gZt.reshape(3 * M, N)]))
                gb = torch.empty((N,), device=X.device, dtype=torch.float32)
                _check(nat.lib().mdl_cv_colsum(_p(gZ), _p(gb), M, N, _stream()))
                return gX, gT, gW, gb, None

        return _F

    _cached = None

    @classmethod
    def apply(cls, X, T, W, b, act):
        if cls._cached is None:
            cls._cached = cls._fn()
        return cls._cached.apply(X, T, W, b, act)


class SteinNet:
    """g: descriptors -> 3 heads x 3-vectors; parameters in FP32 (rounded to
    bf16 at the products). The Stein sums are scaled by ``out_scale`` so that
    outputs of order one give sums of order GPa."""

    def __init__(self, seed=0, device="cuda:0", scale=0.1, out_scale=1e-2):
        torch = _torch()
        gen = torch.Generator().manual_seed(seed)

        def lin(k, n, std):
            return (torch.randn((k, n), generator=gen) * std).to(device).requires_grad_()

        self.W1 = lin(NF, HIDDEN, 1.0 / math.sqrt(NF))
        self.b1 = torch.zeros(HIDDEN, device=device, requires_grad=True)
        self.W2 = lin(HIDDEN, HIDDEN, 1.0 / math.sqrt(HIDDEN))
        self.b2 = torch.zeros(HIDDEN, device=device, requires_grad=True)
        W3 = torch.randn((HIDDEN, NOUT), generator=gen) * (1e-2 / math.sqrt(HIDDEN))
        W3[:, 9:] = 0.0
        self.W3 = W3.to(device).requires_grad_()
        self.b3 = torch.zeros(NOUT, device=device, requires_grad=True)
        self.scale = scale
        self.out_scale = out_scale

    def parameters(self):
        return [self.W1, self.b1, self.W2, self.b2, self.W3, self.b3]

    def numpy_params(self):
        return [p.detach().double().cpu().numpy() for p in self.parameters()]

    def forward(self, D, J):
        """g [M, 9] and its tangents [3, M, 9] (d g / d x_{i,a} of each atom's own a)."""
        X, T = D * self.scale, J * self.scale
        h, ht = TangentDense.apply(X, T, self.W1, self.b1, 0)
        h, ht = TangentDense.apply(h, ht, self.W2, self.b2, 0)
        o, ot = TangentDense.apply(h, ht, self.W3, self.b3, 1)
        return o[:, :9], ot[:, :, :9]

    def stein(self, D, J, F, beta, nframes):
        """s [nframes, 3]: per head out_scale * sum_i (div_i g_i + beta g_i . F_i).
        D [B*N, 64], J [3, B*N, 64], F [B*N, 3] (eV/A), beta in 1/eV."""
        torch = _torch()
        g, gt = self.forward(D, J)
        div = torch.stack([gt[0, :, 3 * c] + gt[1, :, 3 * c + 1] + gt[2, :, 3 * c + 2]
                           for c in range(3)], dim=1)                   # [M, 3]
        gf = torch.stack([(g[:, 3 * c:3 * c + 3] * F).sum(1) for c in range(3)], dim=1)
        per_atom = div + beta * gf
        return self.out_scale * per_atom.reshape(nframes, -1, 3).sum(1)


# ---------------------------------------------------------------------------
# global fields: displacements from the lattice sites, times powers of U
# ---------------------------------------------------------------------------
NGLOBAL = 27


def global_features(u, F, U, U0, beta, ndiv):
    """Stein sums of the fields e_a u_ib, (U - U0) e_a u_ib and (U - U0)^2 e_a u_ib
    of each frame (u [B, N, 3] in A, F [B, N, 3] in eV/A, U [B] in eV):
      S_ab = ndiv d_ab + beta sum_i u_ib F_ia
      T_ab = (U - U0) S_ab - sum_i F_ia u_ib
      Q_ab = (U - U0)^2 S_ab - 2 (U - U0) sum_i F_ia u_ib
    (grad_i U = -F_i). ndiv is the divergence of e_a u_ia on the sampled space:
    N - 1 under dynamics that hold the centre of mass (the fields are tangent
    to that constraint: sum_i m_i u_i = 0 when the sites share the centre of
    mass of the start), N otherwise. Each sum has zero canonical mean for any
    constant U0 while the atoms stay far closer to their sites than half the
    box. Returns [B, 27]."""
    torch = _torch()
    B, N, _ = u.shape
    M = torch.einsum("nia,nib->nab", F, u)
    eye = torch.eye(3, dtype=M.dtype, device=M.device)
    S = ndiv * eye + beta * M
    dU = (U - U0)[:, None, None]
    T = dU * S - M
    Q = dU ** 2 * S - 2.0 * dU * M
    return torch.cat([S.reshape(B, 9), T.reshape(B, 9), Q.reshape(B, 9)], dim=1)


def displacements(pos, sites, box):
    """Minimum-image displacement of every atom from its lattice site (A)."""
    torch = _torch()
    d = pos - sites
    L = torch.as_tensor(box, dtype=d.dtype, device=d.device)
    return d - L * torch.round(d / L)


# ---------------------------------------------------------------------------
# the estimator's per-frame integrands (molecular_dynamics.py:825-1087, NVT)
# ---------------------------------------------------------------------------
HEAD_COMPONENTS = {"C11": ((0, 0), (1, 1), (2, 2)), "C12": ((0, 1), (0, 2), (1, 2)),
                   "C44": ((3, 3), (4, 4), (5, 5))}


def head_integrands(born_stress, born_elast, V, kT, sigma_mean=None):
    """Per-frame y of C11, C12, C44 [n, 3] in the units of the stresses: the
    Born term minus (V/kT)(s_a - <s_a>)(s_b - <s_b>), each averaged over its
    three cubic-equivalent components (molecular_dynamics.py:985-999). The
    kinetic term is a constant in NVT and is added by ``kinetic_terms``."""
    sb = np.asarray(born_stress, dtype=float)           # [n, 6]
    cb = np.asarray(born_elast, dtype=float)            # [n, 6, 6]
    mu = sb.mean(0) if sigma_mean is None else np.asarray(sigma_mean)
    d = sb - mu
    f = V / kT
    out = []
    for idx in HEAD_COMPONENTS.values():
        out.append(sum(cb[:, i, j] - f * d[:, i] * d[:, j] for i, j in idx) / 3.0)
    return np.stack(out, axis=1)


def kinetic_terms(rho_n, kT):
    """Ideal-gas terms of the heads: 2 rho kT for C11, 0 for C12, rho kT for C44."""
    return np.array([2.0 * rho_n * kT, 0.0, rho_n * kT])


def block_bootstrap_se(x, nblocks=64, nresample=2000, seed=0):
    """Standard error of the mean of a correlated series: the means of
    contiguous blocks resampled with replacement."""
    x = np.asarray(x, dtype=float)
    n = len(x) // nblocks * nblocks
    if n < nblocks:
        return float("nan")
    bm = x[:n].reshape(nblocks, -1).mean(1)
    rs = np.random.RandomState(seed)
    idx = rs.randint(0, nblocks, size=(nresample, nblocks))
    return float(bm[idx].mean(1).std())


# ---------------------------------------------------------------------------
# frames from the device loop
# ---------------------------------------------------------------------------
@dataclass
class Frames:
    """Sampled configurations of one NVT run (host arrays, A / eV / GPa)."""
    pos: np.ndarray          # [n, N, 3]
    forces: np.ndarray       # [n, N, 3]
    epot: np.ndarray         # [n] configurational energy
    born_stress: np.ndarray  # [n, 6]
    born_elast: np.ndarray   # [n, 6, 6]
    box: np.ndarray          # [3]
    T: float
    rho_n: float             # atoms per A^3

    @property
    def n(self):
        return len(self.epot)

    @property
    def kT(self):
        return U.k_boltz * self.T / U.ev

    def y(self, sigma_mean=None):
        V = float(np.prod(self.box))
        # stresses in GPa, V/kT in A^3/eV: 1 GPa A^3 / eV = 1/160.2176...
        f = 1e9 * 1e-30 / (U.ev * 1e-7)
        return head_integrands(self.born_stress, self.born_elast, V * f, self.kT, sigma_mean)

    def kinetic(self):
        f = 1e9 * 1e-30 / (U.ev * 1e-7)
        return kinetic_terms(self.rho_n, self.kT) / f


def collect_frames(driver, nframes, stride, T):
    """Advance ``driver`` nframes x stride steps and keep positions, forces,
    energies and the Born stress and elasticity of every stride-th step."""
    N = driver.N
    pos = np.zeros((nframes, N, 3))
    frc = np.zeros((nframes, N, 3))
    ep = np.zeros(nframes)
    sb = np.zeros((nframes, 6))
    cb = np.zeros((nframes, 6, 6))
    for k in range(nframes):
        rec = driver.run(stride, sample_stride=stride)
        r = rec[-1]
        pos[k] = driver.positions_array() / U.angstrom
        frc[k] = driver.forces_array() * U.angstrom / U.ev
        ep[k] = r[nat.S_INDEX["epot_atom"]] * N / U.ev
        sb[k] = r[nat.S_BORN_STRESS:nat.S_BORN_STRESS + 6] / U.GPa
        cb[k] = r[nat.S_BORN_ELAST:nat.S_BORN_ELAST + 36].reshape(6, 6) / U.GPa
    box = np.diag(driver.box.h) / U.angstrom
    return Frames(pos, frc, ep, sb, cb, box, T, N / float(np.prod(box)))


# ---------------------------------------------------------------------------
# the control variate
# ---------------------------------------------------------------------------
class ControlVariate:
    """s_h(x) = theta_h . t(x) + phi_h s_loc,h(x) per head h: t are the 27 global
    Stein sums (scaled by their training spread), s_loc the learned local sum.
    Heads whose held-out variance the variate does not cut are switched off
    (s_h = 0), so the controlled estimate falls back to the plain one."""

    def __init__(self, sites, box, T, device="cuda:0", seed=0, basis: Basis = Basis(),
                 use_global=True, use_net=True):
        torch = _torch()
        self.device = device
        self.sites = torch.as_tensor(np.asarray(sites), dtype=torch.float64, device=device)
        self.box = np.asarray(box, dtype=float)
        self.beta = 1.0 / (U.k_boltz * T / U.ev)
        self.basis = basis
        self.net = SteinNet(seed=seed, device=device)
        self.use_global, self.use_net = use_global, use_net
        self.t_scale = np.ones(NGLOBAL)
        self.U0 = 0.0
        self.coef = np.zeros((3, NGLOBAL + 1))
        self.enabled = np.ones(3, dtype=bool)
        self.y_mean = np.zeros(3)

    # -- per-frame inputs on the device --------------------------------------
    def cache(self, fr: Frames):
        """Keep the descriptors of every frame of ``fr`` on the device (4 MB per
        frame of 4000 atoms) so that training epochs do not recompute them."""
        torch = _torch()
        pos = torch.as_tensor(fr.pos, dtype=torch.float64, device=self.device)
        Ds, Js = [], []
        for b in range(fr.n):
            D, J = descriptors(pos[b], self._types, self.box, self.basis)
            Ds.append(D)
            Js.append(J)
        fr._dev = (torch.stack(Ds), torch.stack(Js))       # [n, N, 64], [n, 3, N, 64]

    def _inputs(self, fr: Frames, sl):
        torch = _torch()
        pos = torch.as_tensor(fr.pos[sl], dtype=torch.float64, device=self.device)
        F64 = torch.as_tensor(fr.forces[sl], dtype=torch.float64, device=self.device)
        B, N, _ = pos.shape
        if getattr(fr, "_dev", None) is not None:
            D = fr._dev[0][sl].reshape(B * N, NF)
            J = fr._dev[1][sl].permute(1, 0, 2, 3).reshape(3, B * N, NF)
        else:
            Ds, Js = [], []
            for b in range(B):
                D, J = descriptors(pos[b], self._types, self.box, self.basis)
                Ds.append(D)
                Js.append(J)
            D = torch.cat(Ds, 0)
            J = torch.cat(Js, 1)
        u = displacements(pos, self.sites, self.box)
        Ut = torch.as_tensor(fr.epot[sl], dtype=torch.float64, device=self.device)
        t = global_features(u, F64, Ut, self.U0, self.beta, N - 1)
        return D, J, F64.reshape(B * N, 3).float(), t

    def set_types(self, types):
        torch = _torch()
        self._types = torch.as_tensor(np.asarray(types, dtype=np.int32), device=self.device)

    def local_sums(self, fr: Frames, sl, grad=False):
        torch = _torch()
        D, J, F, t = self._inputs(fr, sl)
        B = t.shape[0]
        if not self.use_net:
            return torch.zeros((B, 3), device=self.device), t
        with torch.set_grad_enabled(grad):
            s = self.net.stein(D, J, F, self.beta, B)
        return s, t

    def features(self, fr: Frames, batch=8):
        """[n, 3, 28] per head: the scaled global sums and the local sum."""
        out = np.zeros((fr.n, 3, NGLOBAL + 1))
        for k0 in range(0, fr.n, batch):
            sl = slice(k0, min(fr.n, k0 + batch))
            s, t = self.local_sums(fr, sl)
            tt = t.cpu().numpy() / self.t_scale
            if not self.use_global:
                tt[:] = 0.0
            out[sl, :, :NGLOBAL] = tt[:, None, :]
            out[sl, :, NGLOBAL] = s.detach().double().cpu().numpy()
        return out

    def _fit_linear(self, X, yc, ridge=1e-6):
        coef = np.zeros((3, NGLOBAL + 1))
        for h in range(3):
            A = X[:, h, :]
            G = A.T @ A
            G += ridge * np.trace(G) / len(G) * np.eye(len(G))
            coef[h] = np.linalg.solve(G, A.T @ yc[:, h])
        return coef

    def fit(self, train: Frames, valid: Frames, epochs=30, batch=8, lr=2e-3, log=None):
        """Train on ``train``: the global coefficients by least squares, then the
        network on the residual with Adam, then all coefficients again; switch off
        the heads whose variance on ``valid`` does not drop."""
        torch = _torch()
        self.U0 = float(train.epot.mean())
        # spread of the global sums on the training frames
        self.t_scale = np.ones(NGLOBAL)
        t_all = self.features(train)[:, 0, :NGLOBAL]
        self.t_scale = np.maximum(t_all.std(0), 1e-30)
        y = train.y()
        self.y_mean = y.mean(0)
        yc = y - self.y_mean
        X = self.features(train)
        X[:, :, NGLOBAL] = 0.0
        self.coef = self._fit_linear(X, yc)
        w = 1.0 / np.maximum(yc.var(0), 1e-30)
        if self.use_net and epochs > 0:
            opt = torch.optim.Adam(self.net.parameters(), lr=lr)
            lin = torch.as_tensor(self.coef[:, :NGLOBAL], dtype=torch.float64, device=self.device)
            ts = torch.as_tensor(self.t_scale, dtype=torch.float64, device=self.device)
            wt = torch.as_tensor(w, dtype=torch.float64, device=self.device)
            ytc = torch.as_tensor(yc, dtype=torch.float64, device=self.device)
            rs = np.random.RandomState(1234)
            for ep in range(epochs):
                order = rs.permutation(train.n // batch)
                tot = 0.0
                for bi in order:
                    sl = slice(bi * batch, (bi + 1) * batch)
                    s, t = self.local_sums(train, sl, grad=True)
                    tl = (t / ts) if self.use_global else torch.zeros_like(t)
                    r = ytc[sl] - torch.einsum("bq,hq->bh", tl, lin) - s.double()
                    loss = ((r ** 2) * wt).mean()
                    opt.zero_grad()
                    loss.backward()
                    opt.step()
                    tot += float(loss)
                if log:
                    log(f"epoch {ep} loss {tot / len(order):.4f}")
        X = self.features(train)
        self.coef = self._fit_linear(X, yc)
        yv = valid.y()
        sv = self.controls(valid, enabled_only=False)
        var_plain = yv.var(0)
        var_cv = (yv - sv).var(0)
        self.enabled = var_cv < var_plain
        return dict(var_plain_valid=var_plain, var_cv_valid=var_cv, enabled=self.enabled.copy())

    def controls(self, fr: Frames, enabled_only=True):
        """s [n, 3] of each frame (zero for switched-off heads)."""
        X = self.features(fr)
        s = np.einsum("nhq,hq->nh", X, self.coef)
        if enabled_only:
            s = s * self.enabled[None, :]
        return s


def estimate(fr: Frames, s=None, nblocks=64):
    """C11, C12, C44 (GPa) of one run with their block-bootstrap SE, plain or
    controlled by s [n, 3]."""
    y = fr.y()
    if s is not None:
        y = y - s
    C = y.mean(0) + fr.kinetic()
    se = np.array([block_bootstrap_se(y[:, h], nblocks) for h in range(3)])
    return C, se


def subset(fr: Frames, sl) -> Frames:
    return Frames(fr.pos[sl], fr.forces[sl], fr.epot[sl], fr.born_stress[sl], fr.born_elast[sl],
                  fr.box, fr.T, fr.rho_n)


def frames_needed(se, n, target=1.0):
    """Frames for 2 SE <= target (GPa) at the spread seen in n frames."""
    return n * (2.0 * np.asarray(se) / target) ** 2


# ---------------------------------------------------------------------------
# the 4000-atom alpha-brass experiment
# ---------------------------------------------------------------------------
A_300K = 3.6345 * 1.0089030761549767   # A: the zero-pressure lattice constant at 300 K


def brass_driver(edge=10, x_zn=0.06, T=300.0, place_seed=100, vel_seed=200, precision="mixed"):
    """NVT alpha-brass cell on the device loop, and its lattice sites (A) and types."""
    from . import get_potential
    from . import replicas as R
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    spec = dict(edge=edge, N=4 * edge ** 3, x_zn=x_zn, T=T, place_seed=place_seed,
                vel_seed=vel_seed)
    drv = R.make_replica_driver(spec, pots, eams, masses, a=A_300K * U.angstrom,
                                precision=precision)
    sites = R.fcc_positions(edge, A_300K)
    types = np.asarray(drv.types, dtype=np.int32)
    return drv, sites, types


def time_steps(drv, nsteps):
    """Seconds per MD step of the device loop (records off, list rebuilds on schedule)."""
    import time
    drv.run(200, sample_stride=200)
    t0 = time.perf_counter()
    drv.run(nsteps, sample_stride=nsteps)
    return (time.perf_counter() - t0) / nsteps


def time_controls(cvm: ControlVariate, fr: Frames, nframes=32):
    """Device seconds per frame of the controls: descriptors, network and sums."""
    torch = _torch()
    fr = subset(fr, slice(0, nframes))
    fr._dev = None
    cvm.features(fr, batch=8)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    cvm.features(fr, batch=8)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / nframes


def brass_experiment(n_train=2000, n_valid=400, n_eval=4000, eval_runs=2, stride=10,
                     equil=20_000, edge=10, epochs=30, T=300.0, use_net=True, log=None,
                     precision="mixed"):
    """Train the control variate on a plain run and compare controlled and plain
    C11, C12, C44 on fresh trajectories (new velocity seeds, same alloy sample).
    Returns a dict of estimates, SEs, variance ratios and GPU-seconds to
    2 SE <= 1 GPa."""
    torch = _torch()
    say = log or (lambda *_: None)
    drv, sites, types = brass_driver(edge=edge, T=T, vel_seed=200, precision=precision)
    drv.run(equil, sample_stride=equil)
    say(f"training run: {n_train + n_valid} frames every {stride} steps")
    fr = collect_frames(drv, n_train + n_valid, stride, T)
    train, valid = subset(fr, slice(0, n_train)), subset(fr, slice(n_train, None))
    cvm = ControlVariate(sites, fr.box, T, use_net=use_net)
    cvm.set_types(types)
    cvm.cache(train)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    fit = cvm.fit(train, valid, epochs=epochs, log=say)
    e1.record()
    torch.cuda.synchronize()
    train_s = e0.elapsed_time(e1) * 1e-3
    train._dev = None
    t_step = time_steps(drv, 2000)
    t_cv = time_controls(cvm, valid)
    say(f"step {t_step * 1e6:.1f} us, controls {t_cv * 1e6:.1f} us/frame, training {train_s:.1f} s")
    runs = []
    for r in range(eval_runs):
        d, _, _ = brass_driver(edge=edge, T=T, vel_seed=300 + r, precision=precision)
        d.run(equil, sample_stride=equil)
        ev = collect_frames(d, n_eval, stride, T)
        s = cvm.controls(ev)
        Cp, sep = estimate(ev)
        Cc, sec = estimate(ev, s)
        y = ev.y()
        runs.append(dict(C_plain=Cp, se_plain=sep, C_cv=Cc, se_cv=sec,
                         var_ratio=(y - s).var(0) / y.var(0),
                         mean_s=s.mean(0), se_mean_s=np.array(
                             [block_bootstrap_se(s[:, h]) for h in range(3)])))
        say(f"eval run {r}: plain {np.round(Cp, 2)} +- {np.round(sep, 3)}; "
            f"cv {np.round(Cc, 2)} +- {np.round(sec, 3)}")
        del d
    se_p = np.sqrt(np.mean([x["se_plain"] ** 2 for x in runs], 0))
    se_c = np.sqrt(np.mean([x["se_cv"] ** 2 for x in runs], 0))
    cost_p = stride * t_step
    cost_c = stride * t_step + t_cv
    gpu_s_plain = frames_needed(se_p, n_eval) * cost_p
    gpu_s_cv = frames_needed(se_c, n_eval) * cost_c
    out = dict(N=4 * edge ** 3, stride=stride, n_train=n_train, n_eval=n_eval, eval_runs=eval_runs,
               t_step_s=t_step, t_controls_s=t_cv, train_s=train_s, enabled=cvm.enabled.tolist(),
               fit=fit, runs=runs, se_plain=se_p, se_cv=se_c,
               gpu_s_plain=gpu_s_plain, gpu_s_cv=gpu_s_cv,
               gpu_s_plain_worst=float(gpu_s_plain.max()),
               gpu_s_cv_worst=float(gpu_s_cv.max()),
               gpu_s_cv_worst_with_training=float(gpu_s_cv.max() + train_s +
                                                  (n_train + n_valid) * stride * t_step))
    out["speedup"] = out["gpu_s_plain_worst"] / out["gpu_s_cv_worst"]
    return out
