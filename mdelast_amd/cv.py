"""Learned Stein control variate for the NVT stress-fluctuation estimator.

New work (no counterpart in the reference). For the canonical distribution
p(x) ~ exp(-beta U(x)) and any smooth vector field g that is periodic in the
box, E_p[div g - beta g . grad U] = 0 (integrate div(g p) over the torus). So
with F = -grad U the Stein sum

    s[g](x) = sum_i [ div_i g_i(x) + beta g_i(x) . F_i ]

has zero mean, and mean(y - sum_q theta_q s[g_q]) estimates the same C_ij as
the plain mean(y) for any fields g_q and fixed coefficients theta.

The per-frame integrand of C11 or C12 is dominated by (V/kT) ds_a ds_b, a
product of two global stress fluctuations; a sum of local terms cannot
follow it. The fields therefore include products with global scalars H whose
gradients are known: s[H g] = H s[g] + sum_i grad_i H . g_i. The fields are

* the displacements from the lattice sites, e_a u_ib, as they are and times
  U - U0 (grad_i U = -F_i) and h - h0 (below); their sums need only u, F, U,
  h and grad h (``global_features``, ``h_features``);
* learned fields c_i = sum_k w_k(D_i) grad_i D_ik: a per-atom network on the
  radial-basis descriptors D_i (32 Gaussians per neighbour species, two hidden
  tanh layers of width 64) weights the descriptor gradients, which gives
  vectors along the bonds of atom i (equivariant). Its divergence,
  sum_k [w_k L_k + grad_i w_k . grad_i D_k] with L the descriptor Laplacian,
  needs the derivatives of the weights along each atom's own coordinates,
  which the network carries forward. Six fields: per head (C11, C12, C44) one
  used as is and one weighted by h - h0, where h = sum_i V[t_i] . D_i is a
  linear function of the descriptors fitted to the instantaneous pressure
  (grad h follows from the descriptor derivatives alone).

* zero-variance fields of a stress proxy: h_v = sum over ordered pairs of
  psi_{t_i t_j}(r) n_a n_b (one radial function per species pair, 32
  Gaussians, shared by the six Voigt components v = (a, b), fitted to the
  instantaneous Born stress, R^2 ~ 0.99), and the fields grad h_a and
  (h_a - h0_a) grad h_b, whose Stein sums are lap h_a + beta grad h_a . F and
  grad h_a . grad h_b + (h_a - h0_a)(lap h_b + beta grad h_b . F). In a
  harmonic crystal with grad h along the soft directions the latter is
  const - beta k (h_a - h0_a)(h_b - h0_b): it carries the products of stress
  fluctuations that dominate the integrands, deviatoric parts included.

The coefficients of each head are fitted by least squares (after the network,
when it is used, has been trained on the same frames by Adam on the residual
variance); a head whose held-out variance the variate does not cut is switched
off, and that head keeps the plain estimator. ``crossfit_controls`` needs no
training run: each half of a run is controlled by a variate fitted on the
other half.

Device work is hand-written HIP (include/mdelast_cv.h): descriptors with
their derivatives and Laplacians from the driver's neighbour list; the stress
proxy's basis sums and its per-atom gradients and Laplacians (FP64, all pairs);
and the dense layers with their tangents (Y = act(XW + b), Y'_a = act'(XW + b)
T_a W) and backward on bf16 MFMA with fp32 accumulation. PyTorch holds the device
memory, runs the optimiser and the per-atom contractions; ``TangentDense`` is
the torch.autograd.Function over the kernels.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from . import units as U

NB = 32            # radial functions per neighbour species
NF = 2 * NB        # descriptor length (two species)
HIDDEN = 64
NFIELD = 6         # per head: one field as is, one weighted by h - h0
NOUT = NFIELD * NF
HEADS = ("C11", "C12", "C44")
NGLOBAL = 27       # global_features: S, T, Q
NGLOBAL_USED = 27
FIELD_GAIN = 0.05     # least held-out variance cut that keeps the learned fields of a head
FRAME_BATCH = 256     # stored frames per pass of the feature and proxy-fit kernels
PAIR_WIDTH = 192      # pairs kept per atom and frame (a row that needs more is an error)


def _torch():
    import torch
    return torch


# torch and the library must share one HIP runtime: torch first (see
# _native.lib). The plain MD path never imports torch.
if nat._lib is not None and not nat.BOUND_WITH_TORCH:
    raise ImportError(
        "mdelast_amd.cv needs torch, and the library was loaded without it (bound to the "
        "system HIP runtime; torch would map a second runtime with no device). Import "
        "mdelast_amd.cv (or torch) before the first driver or library call of the process.")
try:
    import torch  # noqa: F401,E402  (the runtime the library will bind)
except ImportError:
    pass


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
    """D [N, 64], J [3, N, 64] and L [N, 64] on the device for FP64 positions
    in A (brute force over pairs)."""
    torch = _torch()
    pos = pos_A.contiguous()
    N = pos.shape[0]
    D = torch.empty((N, NF), device=pos.device, dtype=torch.float32)
    J = torch.empty((3, N, NF), device=pos.device, dtype=torch.float32)
    L = torch.empty((N, NF), device=pos.device, dtype=torch.float32)
    box = (C.c_double * 3)(*[float(b) for b in box_A])
    _check(nat.lib().mdl_cv_descriptors(_p(pos), _p(types.contiguous()), N, box,
                                        basis.rc, basis.mu0, basis.dmu, basis.width,
                                        _p(D), _p(J), _p(L), _stream()))
    return D, J, L


NSTRESS = 42       # zero-variance fields of the stress proxy: 6 gradients, 36 weighted
SP_BLOCKS = 64     # blocks per frame of the basis sums


def _pair_args(pairs):
    """(pairs, counts, width) device pointers of captured pair lists, or nulls."""
    if pairs is None:
        return None, None, 0
    nbr, cnt = pairs
    return _p(nbr.contiguous()), _p(cnt.contiguous()), int(nbr.shape[-1])


def stress_basis_sums(pos_A, types32, box_A, basis: Basis = Basis(), pairs=None):
    """Z [B, 3, 32, 6] (FP64, device) of B frames of positions [B, N, 3] in A:
    over ordered pairs of species pair p = t_i + t_j, sum fc g_k n_a n_b
    (mdl_cv_stress_basis_sums); ``pairs`` = (lists [B, N, W], counts [B, N]) as
    capture_frames keeps them, else the pairs are found among all atoms."""
    torch = _torch()
    pos = pos_A.contiguous()
    B, N, _ = pos.shape
    part = torch.empty((B, SP_BLOCKS, 3, NB, 6), dtype=torch.float64, device=pos.device)
    box = (C.c_double * 3)(*[float(b) for b in box_A])
    _check(nat.lib().mdl_cv_stress_basis_sums(_p(pos), _p(types32), N, B, box, basis.rc,
                                              basis.mu0, basis.dmu, basis.width,
                                              *_pair_args(pairs), _p(part), SP_BLOCKS,
                                              _stream()))
    return part.sum(1)


def _frame_pairs(store, sl):
    """The captured pair lists of frames ``sl`` of a store, or None."""
    if store is None or "nbr" not in store:
        return None
    return store["nbr"][sl], store["ncnt"][sl]


def stress_proxy(pos_A, types32, box_A, basis: Basis, coef):
    """For B frames [B, N, 3] (A) and the proxy's coefficients [3, 32]: h's pair
    sum [B, 6], its Laplacian summed over the atoms [B, 6] and its per-atom
    gradients [B, N, 6, 3] (mdl_cv_stress_proxy)."""
    torch = _torch()
    pos = pos_A.contiguous()
    B, N, _ = pos.shape
    grad = torch.empty((B, N, 6, 3), dtype=torch.float64, device=pos.device)
    hl = torch.empty((B, N, 12), dtype=torch.float64, device=pos.device)
    box = (C.c_double * 3)(*[float(b) for b in box_A])
    _check(nat.lib().mdl_cv_stress_proxy(_p(pos), _p(types32), N, B, box, basis.rc, basis.mu0,
                                         basis.dmu, basis.width, _p(coef.contiguous()),
                                         _p(grad), _p(hl), _stream()))
    s = hl.sum(1)
    return s[:, :6], s[:, 6:], grad


SF_BLOCKS = 64     # blocks per frame of the field sums
SF_N = 39          # per-frame field sums
# (a, b), a <= b, row by row: the order of the 21 gradient products
SF_PAIRS = [(a, b) for a in range(6) for b in range(a, 6)]


def stress_field_sums(pos_A, F, types32, box_A, basis: Basis, coef, pairs=None):
    """[B, 39] (FP64, device) for B frames of positions [B, N, 3] in A and
    forces F [B, N, 3] (in the units the sums want): over the atoms m, the
    proxy's pair sum h_v (6), its Laplacian lap_m h_v (6), grad_m h_v . F_m (6)
    and grad_m h_a . grad_m h_b for a <= b (21) (mdl_cv_stress_fields);
    ``pairs`` as for stress_basis_sums."""
    torch = _torch()
    pos = pos_A.contiguous()
    Fc = F.contiguous()
    B, N, _ = pos.shape
    part = torch.empty((B, SF_BLOCKS, SF_N), dtype=torch.float64, device=pos.device)
    box = (C.c_double * 3)(*[float(b) for b in box_A])
    _check(nat.lib().mdl_cv_stress_fields(_p(pos), _p(Fc), 1.0, _p(types32), N, B, box,
                                          basis.rc, basis.mu0, basis.dmu, basis.width,
                                          _p(coef.contiguous()), *_pair_args(pairs), _p(part),
                                          SF_BLOCKS, _stream()))
    return part.sum(1)


def driver_descriptors(drv, basis: Basis = Basis(), device="cuda:0", out=None, La=None):
    """D, J, L of the driver's current configuration from its neighbour list,
    and its positions and forces (CGS, [N, 3]); ``out`` = (D, J, L, pos, F)
    device buffers to fill, else new tensors. With ``La`` (a [3, N, 64] FP32
    device tensor) also the Laplacian's axis terms."""
    torch = _torch()
    N = drv.N
    if out is None:
        out = (torch.empty((N, NF), device=device, dtype=torch.float32),
               torch.empty((3, N, NF), device=device, dtype=torch.float32),
               torch.empty((N, NF), device=device, dtype=torch.float32),
               torch.empty((N, 3), device=device, dtype=torch.float64),
               torch.empty((N, 3), device=device, dtype=torch.float64))
    D, J, L, pos, F = out
    for t in out + ((La,) if La is not None else ()):
        if not t.is_contiguous():
            raise ValueError("driver_descriptors: output buffers must be contiguous")
    # on torch's current stream, ordered against the driver's stream by events
    _check(nat.lib().mdl_driver_cv_descriptors(drv.handle, basis.rc, basis.mu0, basis.dmu,
                                               basis.width, _p(D), _p(J), _p(L), _p(La),
                                               _p(pos), _p(F), _stream()))
    return out


def driver_atoms(drv, what, device="cuda:0"):
    """A per-atom [N, 3] array of the driver (MDL_A_* layout, CGS) copied on the device."""
    torch = _torch()
    out = torch.empty((drv.N, 3), device=device, dtype=torch.float64)
    _check(nat.lib().mdl_driver_copy_atoms(drv.handle, nat.A_INDEX[what], _p(out)))
    return out


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


def _colsum(X):
    torch = _torch()
    X = X.contiguous()
    R, N = X.shape
    work = torch.empty(((R + 1023) // 1024, N), device=X.device, dtype=torch.float32)
    out = torch.empty((N,), device=X.device, dtype=torch.float32)
    _check(nat.lib().mdl_cv_colsum(_p(X), _p(out), _p(work), R, N, _stream()))
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
                if act == 1:            # identity: the products are the outputs
                    ctx.save_for_backward(X, T, W, Z, Zt)
                    ctx.act = act
                    return Z, Zt
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
                if ctx.act == 1:
                    gZ, gZt = gY, gYt
                else:
                    gZ = torch.empty_like(Z)
                    gZt = torch.empty_like(Zt)
                    _check(nat.lib().mdl_cv_act_backward(_p(Z), _p(Zt), _p(gY), _p(gYt),
                                                         _p(gZ), _p(gZt), M * N, ctx.act,
                                                         _stream()))
                gX = _gemm(gZ, W, transB=True)
                gT = _gemm(gZt.reshape(3 * M, N), W, transB=True).reshape(3, M, K)
                gW = _gemm_tn(torch.cat([X, T.reshape(3 * M, K)]),
                              torch.cat([gZ, gZt.reshape(3 * M, N)]))
                return gX, gT, gW, _colsum(gZ), None

        return _F

    _cached = None

    @classmethod
    def apply(cls, X, T, W, b, act):
        if cls._cached is None:
            cls._cached = cls._fn()
        return cls._cached.apply(X, T, W, b, act)


class FieldContract:
    """(a, cF, cG) per atom and field from the field weights w [M, F*64] and
    their tangents wt [3, M, F*64] (mdl_cv_field_contract), with the backward
    with respect to w and wt (mdl_cv_field_contract_backward)."""

    _cached = None

    @staticmethod
    def _fn():
        torch = _torch()

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, w, wt, J, L, F, G, beta):
                M = J.shape[1]
                nf = w.shape[1] // NF
                outs = [torch.empty((M, nf), device=w.device, dtype=torch.float32)
                        for _ in range(3)]
                args = [x.contiguous() for x in (w, wt, J, L, F, G)]
                _check(nat.lib().mdl_cv_field_contract(*[_p(x) for x in args], float(beta), M, nf,
                                                       *[_p(o) for o in outs], _stream()))
                ctx.save_for_backward(*args[2:])
                ctx.beta, ctx.nf = float(beta), nf
                return tuple(outs)

            @staticmethod
            def backward(ctx, ga, gcf, gcg):
                J, L, F, G = ctx.saved_tensors
                M, nf = J.shape[1], ctx.nf
                g = [torch.zeros((M, nf), device=J.device) if x is None else x.contiguous()
                     for x in (ga, gcf, gcg)]
                gw = torch.empty((M, nf * NF), device=J.device)
                gwt = torch.empty((3, M, nf * NF), device=J.device)
                _check(nat.lib().mdl_cv_field_contract_backward(
                    _p(J), _p(L), _p(F), _p(G), ctx.beta, *[_p(x) for x in g], M, nf, _p(gw),
                    _p(gwt), _stream()))
                return gw, gwt, None, None, None, None, None

        return _F

    @classmethod
    def apply(cls, w, wt, J, L, F, G, beta):
        if cls._cached is None:
            cls._cached = cls._fn()
        return cls._cached.apply(w, wt, J, L, F, G, beta)


class SteinNet:
    """Weights of the learned fields: descriptors -> 6 fields x 64 weights per
    atom, with their tangents along each atom's own coordinates. Parameters in
    FP32, rounded to bf16 at the products. Field sums are scaled by
    ``out_scale``."""

    def __init__(self, seed=0, device="cuda:0", scale=0.1, out_scale=1e-2):
        torch = _torch()
        gen = torch.Generator().manual_seed(seed)

        def lin(k, n, std):
            return (torch.randn((k, n), generator=gen) * std).to(device).requires_grad_()

        self.W1 = lin(NF, HIDDEN, 1.0 / math.sqrt(NF))
        self.b1 = torch.zeros(HIDDEN, device=device, requires_grad=True)
        self.W2 = lin(HIDDEN, HIDDEN, 1.0 / math.sqrt(HIDDEN))
        self.b2 = torch.zeros(HIDDEN, device=device, requires_grad=True)
        self.W3 = lin(HIDDEN, NOUT, 1e-2 / math.sqrt(HIDDEN))
        self.b3 = torch.zeros(NOUT, device=device, requires_grad=True)
        self.scale = scale
        self.out_scale = out_scale

    def parameters(self):
        return [self.W1, self.b1, self.W2, self.b2, self.W3, self.b3]

    def numpy_params(self):
        return [p.detach().double().cpu().numpy() for p in self.parameters()]

    def weights(self, D, J):
        """The field weights w [M, 6*64] and their tangents wt [3, M, 6*64]."""
        X, T = D * self.scale, J * self.scale
        h, ht = TangentDense.apply(X, T, self.W1, self.b1, 0)
        h, ht = TangentDense.apply(h, ht, self.W2, self.b2, 0)
        return TangentDense.apply(h, ht, self.W3, self.b3, 1)

    def forward(self, D, J):
        """w [M, 6, 64] and its tangents wt [3, M, 6, 64]."""
        o, ot = self.weights(D, J)
        M = D.shape[0]
        return o.view(M, NFIELD, NF), ot.view(3, M, NFIELD, NF)

    def field_sums(self, D, J, L, F, G, beta, nframes):
        """Per frame and field: A = sum_i div_i c_i + beta c_i . F_i (the Stein
        sum), C = sum_i c_i . F_i and E = sum_i c_i . G_i, all times out_scale
        ([nframes, 6] each), for D [B*N, 64], J [3, B*N, 64], L [B*N, 64],
        F and G [B*N, 3] (F in eV/A), beta in 1/eV."""
        o, ot = self.weights(D, J)
        a, cF, cG = FieldContract.apply(o, ot, J, L, F, G, beta)
        A = a.view(nframes, -1, NFIELD).sum(1)
        Cs = cF.view(nframes, -1, NFIELD).sum(1)
        E = cG.view(nframes, -1, NFIELD).sum(1)
        return self.out_scale * A, self.out_scale * Cs, self.out_scale * E


def net_features(A, Cs, E, dh):
    """[B, 3, 2] per head: the Stein sums of its plain field and of its field
    weighted by h - h0 (dh A + E, E = sum_i c_i . grad_i h)."""
    torch = _torch()
    A, E = A.double(), E.double()
    return torch.stack([A[:, 0::2], dh[:, None] * A[:, 1::2] + E[:, 1::2]], dim=2)


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
    """Minimum-image displacement of every atom from its lattice site (A); box
    a sequence or a device tensor of the three edges."""
    torch = _torch()
    d = pos - sites
    L = box if torch.is_tensor(box) else torch.as_tensor(box, dtype=d.dtype, device=d.device)
    return d - L * torch.round(d / L)


def proxy_weights(V):
    """Gradient weights of h = sum_i V[t_i] . D_i: grad_k h = sum_q Wg[t_k, q] J_k,q
    with Wg[c, t*32 + m] = V[c, t*32 + m] + V[t, c*32 + m] (atom k's own
    descriptors, and its neighbours' channel of k's species)."""
    Wg = V.clone()
    for c in range(2):
        for t in range(2):
            Wg[c, t * NB:(t + 1) * NB] += V[t, c * NB:(c + 1) * NB]
    return Wg


def h_features(S, dh, gh, u):
    """Stein sums of the fields (h - h0) e_a u_ib: dh S_ab + sum_k (grad_k h)_a u_kb, [B, 9]."""
    torch = _torch()
    return (dh[:, None, None] * S + torch.einsum("bna,bnc->bac", gh, u)).reshape(-1, 9)


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
    """Scalars of the sampled steps of one NVT run (A / eV / GPa); with
    ``store`` the per-atom inputs of the controls stay on the device."""
    epot: np.ndarray         # [n] configurational energy
    born_stress: np.ndarray  # [n, 6]
    born_elast: np.ndarray   # [n, 6, 6]
    box: np.ndarray          # [3]
    T: float
    rho_n: float             # atoms per A^3
    store: dict = None       # D [n,N,64], J [n,3,N,64], L [n,N,64], u [n,N,3], F [n,N,3]

    @property
    def n(self):
        return len(self.epot)

    @property
    def kT(self):
        return U.k_boltz * self.T / U.ev

    def y(self, sigma_mean=None):
        V = float(np.prod(self.box))
        # stresses in GPa, V/kT in A^3/eV: 1 GPa A^3 = 1e-21 J
        f = 1e9 * 1e-30 / (U.ev * 1e-7)
        return head_integrands(self.born_stress, self.born_elast, V * f, self.kT, sigma_mean)

    def kinetic(self):
        f = 1e9 * 1e-30 / (U.ev * 1e-7)
        return kinetic_terms(self.rho_n, self.kT) / f


def subset(fr: Frames, sl) -> Frames:
    st = None if fr.store is None else {k: v[sl] for k, v in fr.store.items()}
    return Frames(fr.epot[sl], fr.born_stress[sl], fr.born_elast[sl], fr.box, fr.T, fr.rho_n, st)


def _record_scalars(r, N):
    return (r[nat.S_INDEX["epot_atom"]] * N / U.ev,
            r[nat.S_BORN_STRESS:nat.S_BORN_STRESS + 6] / U.GPa,
            r[nat.S_BORN_ELAST:nat.S_BORN_ELAST + 36].reshape(6, 6) / U.GPa)


def frame_inputs(drv, basis: Basis = Basis(), device="cuda:0", out=None):
    """The per-atom inputs of the controls for the driver's current step, on
    the device: descriptors D, J, L and the positions and forces (CGS).
    ``out`` = (D, J, L, pos, F) buffers to fill."""
    return driver_descriptors(drv, basis, device, out=out)


def collect_frames(drv, nframes, stride, T, sites=None, basis: Basis = Basis(),
                   device="cuda:0"):
    """Advance the driver nframes x stride steps and keep the scalars of every
    stride-th step, and with ``sites`` (a flag: any value but None) its per-atom
    inputs on the device (D, J, L, and positions and forces in CGS)."""
    torch = _torch()
    N = drv.N
    ep = np.zeros(nframes)
    sb = np.zeros((nframes, 6))
    cb = np.zeros((nframes, 6, 6))
    box = np.diag(drv.box.h) / U.angstrom
    store = None
    if sites is not None:
        store = dict(D=torch.empty((nframes, N, NF), device=device),
                     J=torch.empty((nframes, 3, N, NF), device=device),
                     L=torch.empty((nframes, N, NF), device=device),
                     pos=torch.empty((nframes, N, 3), device=device, dtype=torch.float64),
                     F=torch.empty((nframes, N, 3), device=device, dtype=torch.float64))
    for k in range(nframes):
        r = drv.run(stride, sample_stride=stride)[-1]
        ep[k], sb[k], cb[k] = _record_scalars(r, N)
        if store is not None:
            frame_inputs(drv, basis, device,
                         out=tuple(store[q][k] for q in ("D", "J", "L", "pos", "F")))
    return Frames(ep, sb, cb, box, T, N / float(np.prod(box)), store)


# ---------------------------------------------------------------------------
# the control variate
# ---------------------------------------------------------------------------
class ControlVariate:
    """s_h(x) = theta_h . X_h(x) per head h. X_h: 27 global Stein sums (the
    displacement fields e_a u_ib as they are, weighted by U - U0 and weighted by
    h - h0), with ``use_stress`` 42 more from the zero-variance fields of a
    fitted stress proxy (``stress_features``), all scaled by their training
    spread, and the sums of the head's two learned fields (plain, h-weighted).
    h = sum_i V[t_i] . D_i is a linear function of the descriptors fitted to the
    instantaneous pressure. Heads whose held-out variance the variate does not
    cut are switched off (s_h = 0), so the controlled estimate falls back to
    the plain one."""

    def __init__(self, sites, box, T, types, device="cuda:0", seed=0, basis: Basis = Basis(),
                 use_net=True, use_stress=False):
        torch = _torch()
        self.device = device
        self.sites = torch.as_tensor(np.asarray(sites), dtype=torch.float64, device=device)
        self.types = torch.as_tensor(np.asarray(types), dtype=torch.long, device=device)
        self.box = np.asarray(box, dtype=float)
        self.beta = 1.0 / (U.k_boltz * T / U.ev)
        self.basis = basis
        self.net = SteinNet(seed=seed, device=device)
        self.use_net = use_net
        self.use_stress = use_stress
        self.nshared = NGLOBAL_USED + (NSTRESS if use_stress else 0)
        self.NX = self.nshared + 2
        self.types32 = self.types.to(torch.int32).contiguous()
        self.sp_coef = torch.zeros((3, NB), dtype=torch.float64, device=device)
        self.sp_off = np.zeros(6)
        self.sp_h0 = np.zeros(6)
        self.t_scale = np.ones(self.nshared)
        self.U0 = 0.0
        self.h0 = 0.0
        self.V = torch.zeros((2, NF), dtype=torch.float64, device=device)
        self.coef = np.zeros((3, self.NX))
        self.enabled = np.ones(3, dtype=bool)

    # -- the pressure proxy h ----------------------------------------------------
    def _species_sums(self, D):
        """[B, 2, 64]: sum of the descriptors over the atoms of each species."""
        torch = _torch()
        oh = torch.nn.functional.one_hot(self.types, 2).to(D.dtype)      # [N, 2]
        return torch.einsum("bnk,nc->bck", D, oh)

    def fit_proxy(self, fr: Frames, batch=FRAME_BATCH, ridge=1e-8):
        """V by least squares of the pressure -(s_1 + s_2 + s_3)/3 (GPa) on the
        species sums of the descriptors; returns R^2 on ``fr``."""
        torch = _torch()
        P = np.concatenate([self._species_sums(fr.store["D"][k:k + batch]).double().cpu().numpy()
                            for k in range(0, fr.n, batch)]).reshape(fr.n, -1)
        p = -fr.born_stress[:, :3].mean(1)
        Pm, pm = P.mean(0), p.mean()
        A = P - Pm
        G = A.T @ A
        G += ridge * np.trace(G) / len(G) * np.eye(len(G))
        v = np.linalg.solve(G, A.T @ (p - pm))
        self.V = torch.as_tensor(v.reshape(2, NF), dtype=torch.float64, device=self.device)
        self.h0 = float(Pm @ v)
        r = p - pm - A @ v
        return 1.0 - r.var() / p.var()

    def _device_consts(self):
        """The box, the feature scales and the coefficients on the device,
        refreshed when fit() replaces them (no host copies inside the loop)."""
        torch = _torch()
        key = (self.t_scale, self.coef)
        if getattr(self, "_consts_of", None) is None or any(
                a is not b for a, b in zip(self._consts_of, key)):
            dev = self.device
            self._consts = (torch.as_tensor(self.box, dtype=torch.float64, device=dev),
                            torch.as_tensor(self.t_scale, dtype=torch.float64, device=dev),
                            torch.as_tensor(self.coef, dtype=torch.float64, device=dev))
            self._consts_of = key
        return self._consts

    def _proxy_tables(self):
        """Per-atom rows of V (FP64) and of the gradient weights (FP32), cached."""
        if getattr(self, "_tables_of", None) is not self.V:
            self._tables = (self.V[self.types], proxy_weights(self.V)[self.types].float())
            self._tables_of = self.V
        return self._tables

    # -- the stress proxy and its zero-variance fields ---------------------------
    def fit_stress_proxy(self, fr: Frames, batch=FRAME_BATCH, ridge=1e-10):
        """The stress proxy h_v = off_v + sum over ordered pairs of
        psi_{t_i+t_j}(r) n_a n_b (psi = fc sum_k c[p, k] g_k, one radial function
        per species pair shared by the six Voigt components), by least squares
        of the Born stress (GPa) on the basis sums of ``fr``'s stored frames;
        returns R^2 per component."""
        Z = np.concatenate([stress_basis_sums(fr.store["pos"][k:k + batch] / U.angstrom,
                                              self.types32, self.box, self.basis,
                                              _frame_pairs(fr.store, slice(k, k + batch)))
                            .cpu().numpy() for k in range(0, fr.n, batch)])
        n = fr.n
        A = np.concatenate([Z.transpose(0, 3, 1, 2).reshape(n * 6, 3 * NB),
                            np.tile(np.eye(6), (n, 1))], axis=1)
        y = fr.born_stress.reshape(n * 6)
        G = A.T @ A
        G += ridge * np.trace(G) / len(G) * np.eye(len(G))
        c = np.linalg.solve(G, A.T @ y)
        torch = _torch()
        self.sp_coef = torch.as_tensor(c[:3 * NB].reshape(3, NB), dtype=torch.float64,
                                       device=self.device)
        self.sp_off = c[3 * NB:]
        pred = (A @ c).reshape(n, 6)          # h_v of every fitting frame
        self.sp_h0 = pred.mean(0)
        res = fr.born_stress - pred
        return 1.0 - res.var(0) / fr.born_stress.var(0)

    def stress_features(self, pos, F, pairs=None):
        """[B, 42] Stein sums of the zero-variance fields of the stress proxy h
        (positions in CGS, forces in eV/A): the fields grad h_a (sums lap h_a +
        beta sum_m grad_m h_a . F_m) and (h_a - h0_a) grad h_b (sums
        sum_m grad_m h_a . grad_m h_b + (h_a - h0_a) x the former), from the
        per-frame sums of one kernel pass (no per-atom arrays)."""
        torch = _torch()
        s = stress_field_sums(pos / U.angstrom, F, self.types32, self.box, self.basis,
                              self.sp_coef, pairs)
        h = s[:, :6] + torch.as_tensor(self.sp_off, dtype=s.dtype, device=s.device)
        dh = h - torch.as_tensor(self.sp_h0, dtype=s.dtype, device=s.device)
        A = s[:, 6:12] + self.beta * s[:, 12:18]
        GG = torch.empty((s.shape[0], 6, 6), dtype=s.dtype, device=s.device)
        for q, (a, b) in enumerate(SF_PAIRS):
            GG[:, a, b] = s[:, 18 + q]
            GG[:, b, a] = s[:, 18 + q]
        Bm = GG + dh[:, :, None] * A[:, None, :]
        return torch.cat([A, Bm.reshape(-1, 36)], dim=1)

    # -- features ----------------------------------------------------------------
    def _x(self, D, J, L, pos, Fc, Ut, grad=False, pairs=None):
        """X [B, 3, NX] (torch, FP64) for B frames of per-atom inputs (positions
        and forces in CGS)."""
        torch = _torch()
        B, N, _ = pos.shape
        box_d, tsc, _ = self._device_consts()
        u = displacements(pos / U.angstrom, self.sites, box_d)
        F = Fc * (U.angstrom / U.ev)
        M = (F.unsqueeze(3) * u.unsqueeze(2)).sum(1)                        # [B, 3, 3]
        S = (N - 1) * torch.eye(3, dtype=M.dtype, device=M.device) + self.beta * M
        dU = Ut - self.U0
        Vt, Wg = self._proxy_tables()
        dh = (D.double() * Vt).sum((1, 2)) - self.h0
        gh = (J * Wg).sum(3).transpose(1, 2).double()                       # [B, N, 3]
        Th = dh[:, None, None] * S + (gh.unsqueeze(3) * u.unsqueeze(2)).sum(1)
        t = torch.cat([S.reshape(B, 9), (dU[:, None, None] * S - M).reshape(B, 9),
                       Th.reshape(B, 9)], dim=1)
        if self.use_stress:
            t = torch.cat([t, self.stress_features(pos, F, pairs)], dim=1)
        t = t / tsc
        if self.use_net:
            with torch.set_grad_enabled(grad):
                A, Cs, E = self.net.field_sums(
                    D.reshape(B * N, NF), J.permute(1, 0, 2, 3).reshape(3, B * N, NF),
                    L.reshape(B * N, NF), F.reshape(B * N, 3).float(),
                    gh.reshape(B * N, 3).float(), self.beta, B)
                nx = net_features(A, Cs, E, dh)
        else:
            nx = torch.zeros((B, 3, 2), dtype=t.dtype, device=t.device)
        return torch.cat([t[:, None, :].expand(B, 3, self.nshared), nx], dim=2)

    def _batch(self, fr: Frames, sl, grad=False):
        torch = _torch()
        st = fr.store
        Ut = torch.as_tensor(fr.epot[sl], dtype=torch.float64, device=self.device)
        return self._x(st["D"][sl], st["J"][sl], st["L"][sl], st["pos"][sl], st["F"][sl], Ut,
                       grad, _frame_pairs(st, sl))

    def features(self, fr: Frames, batch=FRAME_BATCH):
        out = np.zeros((fr.n, 3, self.NX))
        for k0 in range(0, fr.n, batch):
            sl = slice(k0, min(fr.n, k0 + batch))
            out[sl] = self._batch(fr, sl).detach().cpu().numpy()
        return out

    @staticmethod
    def _fit_linear(X, yc, ridge=1e-6):
        coef = np.zeros((3, X.shape[2]))
        for h in range(3):
            A = X[:, h, :]
            G = A.T @ A
            G += ridge * max(np.trace(G), 1e-300) / len(G) * np.eye(len(G))
            coef[h] = np.linalg.solve(G, A.T @ yc[:, h])
        return coef

    def fit(self, train: Frames, valid: Frames, epochs=30, batch=8, lr=1e-3, weight_decay=1e-4,
            stop_fraction=0.2, log=None):
        """Fit the control variate on ``train`` and decide on ``valid``.

        1. the pressure proxy h (least squares);
        2. the global coefficients (least squares, 27 sums per head);
        3. the network (Adam on the residual variance of the heads that step 2
           helps on ``valid``) on the first part of ``train``, the parameters of
           the epoch with the lowest residual on the last ``stop_fraction`` kept;
        4. all coefficients again (least squares on ``train``);
        5. per head, the controls with the learned fields are kept where they
           beat the global ones on ``valid``, the global ones where those beat
           plain, and the head is switched off otherwise."""
        torch = _torch()
        say = log or (lambda *_: None)
        self.U0 = float(train.epot.mean())
        r2 = self.fit_proxy(train)
        say(f"pressure proxy: R^2 {r2:.4f} on the training frames")
        if self.use_stress:
            r2s = self.fit_stress_proxy(train)
            say(f"stress proxy: R^2 {np.round(r2s, 4)} on the training frames")
        self.t_scale = np.ones(self.nshared)
        use_net, self.use_net = self.use_net, False
        X = self.features(train)
        self.t_scale = np.maximum(X[:, 0, :self.nshared].std(0), 1e-30)
        X[:, :, :self.nshared] /= self.t_scale
        y = train.y()
        yc = y - y.mean(0)
        glob = np.concatenate([self._fit_linear(X[:, :, :self.nshared], yc), np.zeros((3, 2))],
                              axis=1)
        self.coef = glob
        yv = valid.y()
        var_plain = yv.var(0)
        var_glob = (yv - self.controls(valid, enabled_only=False)).var(0)
        say(f"global fields: held-out variance / plain {np.round(var_glob / var_plain, 4)}")
        self.use_net = use_net
        history, var_net = [], np.full(3, np.inf)
        heads = var_glob < var_plain
        if self.use_net and epochs > 0 and heads.any():
            n_fit = int(round(train.n * (1.0 - stop_fraction)))
            fit_part, stop_part = subset(train, slice(0, n_fit)), subset(train, slice(n_fit, None))
            self.coef = np.concatenate([glob[:, :self.nshared], np.ones((3, 2))], axis=1)
            params = self.net.parameters()
            opt = torch.optim.AdamW(params, lr=lr, weight_decay=weight_decay)
            sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=0.1 ** (1.0 / epochs))
            co = torch.as_tensor(self.coef, dtype=torch.float64, device=self.device)
            w = heads / np.maximum(yc.var(0), 1e-30)
            wt = torch.as_tensor(w / w.sum(), dtype=torch.float64, device=self.device)
            ytc = torch.as_tensor(yc, dtype=torch.float64, device=self.device)
            ys = yc[n_fit:]
            rs = np.random.RandomState(1234)
            best, best_state = np.inf, [p.detach().clone() for p in params]
            for ep in range(epochs):
                tot = 0.0
                order = rs.permutation(fit_part.n // batch)
                for bi in order:
                    sl = slice(bi * batch, (bi + 1) * batch)
                    Xb = self._batch(fit_part, sl, grad=True)
                    r = ytc[sl] - (Xb * co[None]).sum(2)
                    loss = ((r ** 2) * wt).sum(1).mean()
                    opt.zero_grad()
                    loss.backward()
                    torch.nn.utils.clip_grad_norm_(params, 1.0)
                    opt.step()
                    tot += float(loss.detach())
                sched.step()
                stop = float(((((ys - self.controls(stop_part, enabled_only=False)) ** 2)
                               * w / w.sum()).sum(1)).mean())
                history.append(dict(fit=tot / len(order), stop=stop))
                if stop < best:
                    best = stop
                    best_state = [p.detach().clone() for p in params]
                say(f"epoch {ep}: residual variance / plain {tot / len(order):.4f} "
                    f"(held-back {stop:.4f})")
            with torch.no_grad():
                for p, q in zip(params, best_state):
                    p.copy_(q)
            X = self.features(train)
            self.coef = self._fit_linear(X, yc)
            var_net = (yv - self.controls(valid, enabled_only=False)).var(0)
            say(f"learned fields: held-out variance / plain {np.round(var_net / var_plain, 4)}")
        # The learned fields cost a network pass per frame, shared by all heads,
        # and the head with the largest controlled/plain variance ratio sets the
        # cost of reaching the target on all three constants. A head keeps its
        # fields where they cut its held-out variance by FIELD_GAIN or more, and
        # the network stays only if that cuts the worst ratio by FIELD_GAIN too.
        use_fields = var_net < (1.0 - FIELD_GAIN) * var_glob
        on_glob = var_glob < var_plain
        on_net = np.where(use_fields, var_net, var_glob) < var_plain

        def worst(v, on):
            r = np.where(on, v / np.maximum(var_plain, 1e-300), 1.0)
            return float(r.max())

        if worst(np.where(use_fields, var_net, var_glob), on_net) >= \
                (1.0 - FIELD_GAIN) * worst(var_glob, on_glob):
            use_fields[:] = False
        coef = self.coef.copy()
        for h in range(3):
            if not use_fields[h]:
                coef[h] = glob[h]
        self.coef = coef          # a new array: the device copy is refreshed
        var_cv = np.where(use_fields, var_net, var_glob)
        self.enabled = var_cv < var_plain
        self.uses_fields = use_fields & self.enabled
        # no head takes the fields: the controls skip the network
        self.use_net = bool(self.uses_fields.any())
        return dict(proxy_r2=r2, var_plain_valid=var_plain, var_global_valid=var_glob,
                    var_fields_valid=var_net, var_cv_valid=var_cv,
                    enabled=self.enabled.copy(), uses_learned_fields=self.uses_fields.copy(),
                    history=history)

    def controls(self, fr: Frames, enabled_only=True):
        """s [n, 3] of stored frames (zero for switched-off heads)."""
        s = np.einsum("nhq,hq->nh", self.features(fr), self.coef)
        return s * self.enabled[None, :] if enabled_only else s


class ControlBatch:
    """Controls of a running driver, gathered frame by frame into device
    buffers and evaluated B frames at a time (one network pass per batch)."""

    def __init__(self, cvm: ControlVariate, N, B=16, pair_width=PAIR_WIDTH):
        torch = _torch()
        dev = cvm.device
        self.cvm, self.B, self.k = cvm, B, 0
        self.buf = (torch.empty((B, N, NF), device=dev), torch.empty((B, 3, N, NF), device=dev),
                    torch.empty((B, N, NF), device=dev),
                    torch.empty((B, N, 3), device=dev, dtype=torch.float64),
                    torch.empty((B, N, 3), device=dev, dtype=torch.float64))
        # pair lists captured by run() (the stress proxy's kernels take them)
        self.width = pair_width
        self.pairs = ((torch.empty((B, N, pair_width), device=dev, dtype=torch.int32),
                       torch.empty((B, N), device=dev, dtype=torch.int32))
                      if pair_width and cvm.use_stress else None)
        self.U = np.zeros(B)
        self.out = []
        # the controls run on the stream current now; the per-frame call takes
        # prepared pointers (slot k of each contiguous buffer) instead of views
        self.stream = torch.cuda.current_stream(dev)
        self._st = C.c_void_p(self.stream.cuda_stream)
        self._slots = [tuple(C.c_void_p(b[k].data_ptr()) for b in self.buf) for k in range(B)]
        bs = cvm.basis
        self._basis = (bs.rc, bs.mu0, bs.dmu, bs.width)
        self._fn = nat.lib().mdl_driver_cv_descriptors

    def push(self, drv, epot):
        """Take the driver's current step (epot: configurational energy, eV)."""
        D, J, L, pos, F = self._slots[self.k]
        _check(self._fn(drv.handle, *self._basis, D, J, L, None, pos, F, self._st))
        self.U[self.k] = epot
        self.k += 1
        if self.k == self.B:
            self.flush()

    def flush(self):
        """Queue the controls of the gathered frames (no host synchronisation)."""
        torch = _torch()
        if self.k == 0:
            return
        k = self.k
        with torch.cuda.stream(self.stream):
            # a fresh pinned copy per batch: the asynchronous upload reads it later
            Ut = torch.tensor(self.U[:k], dtype=torch.float64).pin_memory().to(
                self.cvm.device, non_blocking=True)
            X = self.cvm._x(*[b[:k] for b in self.buf], Ut)
            co = self.cvm._device_consts()[2]
            self.out.append((X * co[None]).sum(2))
        self.k = 0

    def run(self, drv, nframes, stride):
        """Advance ``drv`` by nframes x stride steps, a frame every ``stride``
        steps, and take every frame's controls: the driver captures each
        frame's inputs inside its run (one run per B frames,
        mdl_driver_set_cv_capture) and the controls of those frames are queued
        on the driver's stream right after the run, before the next one
        overwrites the slots. With the stress proxy on, the run also captures
        each frame's pair lists (mdl_driver_set_cv_capture_pairs) for its
        kernels. Returns the frames' records (MDL_S_* rows)."""
        torch = _torch()
        lib = nat.lib()
        ptr = C.c_void_p()
        _check(lib.mdl_driver_stream(drv.handle, C.byref(ptr)))
        self._ext = torch.cuda.ExternalStream(ptr.value, device=self.cvm.device)
        slots = [C.c_void_p(b.data_ptr()) for b in self.buf]
        recs = []
        try:
            for f0 in range(0, nframes, self.B):
                nb = min(self.B, nframes - f0)
                _check(lib.mdl_driver_set_cv_capture(drv.handle, *self._basis, *slots, self.B))
                pairs = None
                if self.pairs is not None:
                    _check(lib.mdl_driver_set_cv_capture_pairs(
                        drv.handle, _p(self.pairs[0]), _p(self.pairs[1]), self.width))
                rec = drv.run(nb * stride, sample_stride=stride)
                recs.append(rec)
                U = np.array([_record_scalars(r, drv.N)[0] for r in rec])
                with torch.cuda.stream(self._ext):
                    if self.pairs is not None:
                        most = int(self.pairs[1][:nb].max())
                        if most > self.width:
                            raise RuntimeError(f"ControlBatch.run: an atom has {most} pairs "
                                               f"within the cut-off, beyond the pair width "
                                               f"{self.width}")
                        pairs = (self.pairs[0][:nb], self.pairs[1][:nb])
                    Ut = torch.tensor(U, dtype=torch.float64).pin_memory().to(
                        self.cvm.device, non_blocking=True)
                    X = self.cvm._x(*[b[:nb] for b in self.buf], Ut, pairs=pairs)
                    co = self.cvm._device_consts()[2]
                    self.out.append((X * co[None]).sum(2))
        finally:
            lib.mdl_driver_set_cv_capture(drv.handle, 0.0, 0.0, 0.0, 0.0, None, None, None,
                                          None, None, 0)
        return np.concatenate(recs) if recs else np.zeros((0, nat.NRECORD))

    def result(self):
        """s [n, 3] of every frame taken so far (all heads)."""
        torch = _torch()
        self.flush()
        if getattr(self, "_ext", None) is not None:
            self.stream.wait_stream(self._ext)
        with torch.cuda.stream(self.stream):
            return torch.cat(self.out).cpu().numpy() if self.out else np.zeros((0, 3))


def estimate(fr: Frames, s=None, nblocks=64):
    """C11, C12, C44 (GPa) of one run with their block-bootstrap SE, plain or
    controlled by s [n, 3]."""
    y = fr.y()
    if s is not None:
        y = y - s
    C = y.mean(0) + fr.kinetic()
    se = np.array([block_bootstrap_se(y[:, h], nblocks) for h in range(3)])
    return C, se


def frames_needed(se, n, target=1.0):
    """Frames for 2 SE <= target (GPa) at the spread seen in n frames."""
    return n * (2.0 * np.asarray(se) / target) ** 2


def capture_frames(drv, nframes, stride, T, basis: Basis = Basis(), device="cuda:0",
                   pair_width=PAIR_WIDTH):
    """One run of nframes x stride steps (one host call) that keeps every
    stride-th step: its scalars from the run's records, and its per-atom inputs
    of the controls (D, J, L, positions, forces) captured on the device inside
    the run (mdl_driver_set_cv_capture with one slot per frame); with
    ``pair_width`` also each atom's candidate pairs, a list in the driver's order
    that holds all its pairs within the cut-off (mdl_driver_set_cv_capture_pairs),
    which the stress proxy's kernels then take, testing r, instead of searching
    all atoms."""
    torch = _torch()
    N = drv.N
    store = dict(D=torch.empty((nframes, N, NF), device=device),
                 J=torch.empty((nframes, 3, N, NF), device=device),
                 L=torch.empty((nframes, N, NF), device=device),
                 pos=torch.empty((nframes, N, 3), device=device, dtype=torch.float64),
                 F=torch.empty((nframes, N, 3), device=device, dtype=torch.float64))
    if pair_width:
        store["nbr"] = torch.empty((nframes, N, pair_width), device=device, dtype=torch.int32)
        store["ncnt"] = torch.empty((nframes, N), device=device, dtype=torch.int32)
    lib = nat.lib()
    _check(lib.mdl_driver_set_cv_capture(drv.handle, basis.rc, basis.mu0, basis.dmu,
                                         basis.width, *[_p(store[q]) for q in
                                                        ("D", "J", "L", "pos", "F")], nframes))
    try:
        if pair_width:
            _check(lib.mdl_driver_set_cv_capture_pairs(drv.handle, _p(store["nbr"]),
                                                       _p(store["ncnt"]), pair_width))
        rec = drv.run(nframes * stride, sample_stride=stride)
    finally:
        lib.mdl_driver_set_cv_capture(drv.handle, 0.0, 0.0, 0.0, 0.0, None, None, None,
                                      None, None, 0)
    if len(rec) != nframes:
        raise RuntimeError(f"capture_frames: {len(rec)} records for {nframes} frames")
    if pair_width:
        most = int(store["ncnt"].max())
        if most > pair_width:
            raise RuntimeError(f"capture_frames: an atom has {most} pairs within the cut-off, "
                               f"beyond the pair width {pair_width}")
    ep = np.zeros(nframes)
    sb = np.zeros((nframes, 6))
    cb = np.zeros((nframes, 6, 6))
    for k in range(nframes):
        ep[k], sb[k], cb[k] = _record_scalars(rec[k], N)
    box = np.diag(drv.box.h) / U.angstrom
    return Frames(ep, sb, cb, box, T, N / float(np.prod(box)), store)


def crossfit_controls(fr: Frames, sites, types, T, folds=2, valid_fraction=0.2,
                      use_net=False, use_stress=True, epochs=0, device="cuda:0", log=None):
    """Controls s [n, 3] of one run's stored frames by cross-fitting, with no
    training run of its own: the run is cut into ``folds`` contiguous blocks, and
    the controls of each block come from a variate fitted (proxy, coefficients,
    head switches; the last ``valid_fraction`` of the fitting frames decides the
    switches) on the other blocks only. Each block's coefficients are fixed with
    respect to its own frames, so its controls keep their zero mean (up to the
    correlation across the block boundary, a few frames of a long run). Returns s
    and each fold's fit summary."""
    n = fr.n
    if folds != 2:
        raise ValueError("crossfit_controls: two folds (each fits on the other half)")
    half = n // 2
    blocks = [slice(0, half), slice(half, n)]
    s = np.zeros((n, 3))
    fits = []
    for k in range(2):
        fit_on = subset(fr, blocks[1 - k])
        m = fit_on.n
        cut = int(round(m * (1.0 - valid_fraction)))
        cvm = ControlVariate(sites, fr.box, T, types, device=device, use_net=use_net,
                             use_stress=use_stress)
        info = cvm.fit(subset(fit_on, slice(0, cut)), subset(fit_on, slice(cut, None)),
                       epochs=epochs, log=log)
        s[blocks[k]] = cvm.controls(subset(fr, blocks[k]))
        fits.append(dict(enabled=cvm.enabled.copy(), uses_learned_fields=cvm.uses_fields.copy(),
                         proxy_r2=info["proxy_r2"],
                         var_ratio_valid=np.asarray(info["var_cv_valid"]) /
                         np.asarray(info["var_plain_valid"])))
    return s, fits


# ---------------------------------------------------------------------------
# the 4000-atom alpha-brass case
# ---------------------------------------------------------------------------
A_300K = 3.6345 * 1.0089030761549767   # A: the zero-pressure lattice constant at 300 K


def brass_driver(edge=10, x_zn=0.06, T=300.0, place_seed=100, vel_seed=200, precision="mixed"):
    """NVT alpha-brass cell on the device loop, its lattice sites (A) and types."""
    from . import get_potential
    from . import replicas as R
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    spec = dict(edge=edge, N=4 * edge ** 3, x_zn=x_zn, T=T, place_seed=place_seed,
                vel_seed=vel_seed)
    drv = R.make_replica_driver(spec, pots, eams, masses, a=A_300K * U.angstrom,
                                precision=precision)
    sites = R.fcc_positions(edge, A_300K)
    return drv, sites, np.asarray(drv.types, dtype=np.int32)


def time_steps(drv, nsteps):
    """Seconds per MD step of the device loop (no records)."""
    drv.run(200, sample_stride=200)
    t0 = time.perf_counter()
    drv.run(nsteps, sample_stride=nsteps)
    return (time.perf_counter() - t0) / nsteps


def frame_loop_time(drv, nframes, stride, cbat=None):
    """Seconds per frame of the frame loop: runs of 16 frames (stride steps
    each, records of every frame), and with ``cbat`` the controls of every
    frame taken as the evaluation takes them (ControlBatch.run)."""
    torch = _torch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if cbat is None:
        for f0 in range(0, nframes, 16):
            drv.run(min(16, nframes - f0) * stride, sample_stride=stride)
    else:
        cbat.run(drv, nframes, stride)
        cbat.result()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / nframes


def brass_experiment(n_train=2000, n_valid=400, n_eval=4000, eval_runs=2, stride=20,
                     equil=20_000, edge=10, epochs=20, T=300.0, use_net=True, log=None,
                     precision="mixed"):
    """Train the control variate on a plain run and compare controlled and plain
    C11, C12, C44 on fresh trajectories (new velocity seeds, same alloy
    sample). The controls of the evaluation runs are computed in the loop from
    the driver's device state, as a production run would. Returns estimates,
    SEs, variance ratios and GPU-seconds to 2 SE <= 1 GPa."""
    torch = _torch()
    say = log or (lambda *_: None)
    drv, sites, types = brass_driver(edge=edge, T=T, vel_seed=200, precision=precision)
    drv.run(equil, sample_stride=equil)
    say(f"training run: {n_train + n_valid} frames every {stride} steps")
    fr = collect_frames(drv, n_train + n_valid, stride, T, sites=sites)
    train, valid = subset(fr, slice(0, n_train)), subset(fr, slice(n_train, None))
    cvm = ControlVariate(sites, fr.box, T, types, use_net=use_net)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fit = cvm.fit(train, valid, epochs=epochs, log=say)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del fr, train, valid
    t_step = time_steps(drv, 4000)
    # marginal device cost of the controls: the same per-frame loop with and
    # without them (the controls run on another stream and overlap the steps)
    loop_plain = frame_loop_time(drv, 400, stride)
    loop_cv = frame_loop_time(drv, 400, stride, ControlBatch(cvm, drv.N))
    t_cv = max(loop_cv - loop_plain, 0.0)
    say(f"step {t_step * 1e6:.1f} us; frame loop {loop_plain * 1e6:.0f} us plain, "
        f"{loop_cv * 1e6:.0f} us with controls; training {train_s:.1f} s; "
        f"enabled {cvm.enabled}")
    runs = []
    for r in range(eval_runs):
        d, _, _ = brass_driver(edge=edge, T=T, vel_seed=300 + r, precision=precision)
        d.run(equil, sample_stride=equil)
        ep = np.zeros(n_eval)
        sb = np.zeros((n_eval, 6))
        cb = np.zeros((n_eval, 6, 6))
        cbat = ControlBatch(cvm, d.N)
        recs = cbat.run(d, n_eval, stride)
        for k in range(n_eval):
            ep[k], sb[k], cb[k] = _record_scalars(recs[k], d.N)
        s = cbat.result()
        box = np.diag(d.box.h) / U.angstrom
        ev = Frames(ep, sb, cb, box, T, d.N / np.prod(box))
        sc = s * cvm.enabled[None, :]
        Cp, sep = estimate(ev)
        Cc, sec = estimate(ev, sc)
        y = ev.y()
        runs.append(dict(C_plain=Cp, se_plain=sep, C_cv=Cc, se_cv=sec,
                         var_ratio_all_heads=(y - s).var(0) / y.var(0),
                         var_ratio=(y - sc).var(0) / y.var(0),
                         mean_s=s.mean(0),
                         se_mean_s=np.array([block_bootstrap_se(s[:, h]) for h in range(3)])))
        say(f"eval run {r}: plain {np.round(Cp, 2)} +- {np.round(sep, 3)}; "
            f"controlled {np.round(Cc, 2)} +- {np.round(sec, 3)}")
        del d, cbat
    se_p = np.sqrt(np.mean([x["se_plain"] ** 2 for x in runs], 0))
    se_c = np.sqrt(np.mean([x["se_cv"] ** 2 for x in runs], 0))
    cost_p = stride * t_step
    cost_c = stride * t_step + t_cv
    gpu_s_plain = frames_needed(se_p, n_eval) * cost_p
    gpu_s_cv = frames_needed(se_c, n_eval) * cost_c
    out = dict(N=4 * edge ** 3, stride=stride, n_train=n_train, n_valid=n_valid, n_eval=n_eval,
               eval_runs=eval_runs, t_step_s=t_step, t_controls_s=t_cv, train_s=train_s,
               frame_loop_s=dict(plain=loop_plain, controlled=loop_cv),
               enabled=cvm.enabled.tolist(), fit=fit, runs=runs, se_plain=se_p, se_cv=se_c,
               gpu_s_plain=gpu_s_plain, gpu_s_cv=gpu_s_cv,
               gpu_s_plain_worst=float(gpu_s_plain.max()),
               gpu_s_cv_worst=float(gpu_s_cv.max()),
               training_gpu_s=float(train_s + (n_train + n_valid) * stride * t_step))
    out["speedup"] = out["gpu_s_plain_worst"] / out["gpu_s_cv_worst"]
    return out


def brass_crossfit_study(n_frames=3000, state_seeds=(101, 102, 103), runs_per_state=3,
                         stride=20, equil=20_000, edge=10, T=300.0, nblocks=32,
                         use_stress=True, precision="mixed", log=None):
    """The control variate as a production run would use it, with no training
    run: on fresh alloy samples (placement seeds ``state_seeds``, none of them
    the sample any default elsewhere uses) and fresh trajectories (velocity
    seeds), each run's frames give the plain estimate and the cross-fitted
    controlled estimate (``crossfit_controls``) of C11, C12, C44.

    Device time per frame: the plain route pays stride MD steps (timed without
    records); the controlled route pays the same steps with the control inputs
    captured inside the run, plus both folds' fits and controls (all timed, so
    the fitting is counted). GPU-s to 2 SE <= 1 GPa = frames needed x time per
    frame, with the pooled SE over all runs (block bootstrap, ``nblocks``
    blocks per run). The control means over the runs, in units of their own
    SE, test the zero-mean property and the SE estimate together."""
    torch = _torch()
    say = log or (lambda *_: None)
    runs = []
    t_step = None
    for p in state_seeds:
        for r in range(runs_per_state):
            d, sites, types = brass_driver(edge=edge, T=T, place_seed=p,
                                           vel_seed=1000 * p + r, precision=precision)
            d.run(equil, sample_stride=equil)
            if t_step is None:
                t_step = time_steps(d, 4000)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fr = capture_frames(d, n_frames, stride, T)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            s, fits = crossfit_controls(fr, sites, types, T, use_stress=use_stress)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            Cp, sep = estimate(fr, nblocks=nblocks)
            Cc, sec = estimate(fr, s, nblocks=nblocks)
            y = fr.y()
            se_s = np.array([block_bootstrap_se(s[:, h], nblocks) for h in range(3)])
            runs.append(dict(place_seed=p, vel_seed=1000 * p + r, C_plain=Cp, se_plain=sep,
                             C_cv=Cc, se_cv=sec, var_ratio=(y - s).var(0) / y.var(0),
                             mean_s=s.mean(0), se_mean_s=se_s,
                             t_capture_frame=(t1 - t0) / n_frames,
                             t_fit_frame=(t2 - t1) / n_frames,
                             enabled=[f["enabled"].tolist() for f in fits]))
            say(f"sample {p} run {r}: plain {np.round(Cp, 2)} +- {np.round(sep, 3)}; "
                f"controlled {np.round(Cc, 2)} +- {np.round(sec, 3)}; "
                f"capture {1e6 * (t1 - t0) / n_frames:.0f} us/frame, fits "
                f"{1e6 * (t2 - t1) / n_frames:.0f} us/frame")
            del fr, d
    se_p = np.sqrt(np.mean([x["se_plain"] ** 2 for x in runs], 0))
    se_c = np.sqrt(np.mean([x["se_cv"] ** 2 for x in runs], 0))
    cost_p = stride * t_step
    cost_c = float(np.mean([x["t_capture_frame"] + x["t_fit_frame"] for x in runs]))
    gpu_s_plain = frames_needed(se_p, n_frames) * cost_p
    gpu_s_cv = frames_needed(se_c, n_frames) * cost_c
    z = np.array([x["mean_s"] / x["se_mean_s"] for x in runs])
    out = dict(N=4 * edge ** 3, stride=stride, n_frames=n_frames, runs=runs,
               state_seeds=list(state_seeds), runs_per_state=runs_per_state, nblocks=nblocks,
               t_step_s=t_step, cost_plain_frame_s=cost_p, cost_cv_frame_s=cost_c,
               se_plain=se_p, se_cv=se_c, gpu_s_plain=gpu_s_plain, gpu_s_cv=gpu_s_cv,
               gpu_s_plain_worst=float(gpu_s_plain.max()), gpu_s_cv_worst=float(gpu_s_cv.max()),
               control_z=z, control_z_rms=np.sqrt((z ** 2).mean(0)))
    out["speedup"] = out["gpu_s_plain_worst"] / out["gpu_s_cv_worst"]
    return out


def profile_case(nframes=32, epochs=1, stride=20, equil=400, log=print):
    """A short pass over every control-variate kernel, for profilers (counter
    runs serialise each dispatch, so the full experiment is too long there):
    the 4,000-atom brass cell, nframes stored training frames, the fit (stress
    proxy included) with ``epochs`` epochs of the network, and the in-loop
    controls of nframes more frames with the network forced on."""
    torch = _torch()
    drv, sites, types = brass_driver()
    drv.run(equil, sample_stride=equil)
    fr = collect_frames(drv, nframes + nframes // 4, stride, 300.0, sites=sites)
    train, valid = subset(fr, slice(0, nframes)), subset(fr, slice(nframes, None))
    cvm = ControlVariate(sites, fr.box, 300.0, types, use_stress=True)
    fit = cvm.fit(train, valid, epochs=epochs)
    cvm.use_net = True            # the controls pass through the network too
    cbat = ControlBatch(cvm, drv.N)
    cbat.run(drv, nframes, stride)
    s = cbat.result()
    torch.cuda.synchronize()
    log(f"profile case: {nframes} frames, {epochs} epoch(s), held-out variance / plain "
        f"{np.round(np.asarray(fit['var_cv_valid']) / np.asarray(fit['var_plain_valid']), 3)}, "
        f"controls of {len(s)} frames")


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="control-variate kernels on the 4,000-atom brass "
                                             "cell (a short case for profilers)")
    ap.add_argument("--frames", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=1)
    a = ap.parse_args()
    profile_case(nframes=a.frames, epochs=a.epochs)
