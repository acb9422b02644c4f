"""CPU restatement of the learned Stein control variate (test infrastructure).

Only the tests and the smoke check import this module: it is the checker of
the device kernels behind mdelast_amd/cv.py (include/mdelast_cv.h), never part
of the product path. The control variate has no counterpart in the reference;
what pins this restatement is mathematics, checked in tests/test_cv_oracle.py:

* descriptor derivatives against central finite differences of the
  descriptors (FP64);
* the descriptor Laplacians against central differences of the derivatives;
* the forward-mode tangents of the network, and the divergence of the fields
  it weights, against finite differences along each atom's own coordinates;
* the Stein identity E[s] = 0 by exact Monte Carlo on a harmonic system whose
  canonical distribution is Gaussian and sampled directly, for the learned
  local field and for every global field (a sign error in the beta term, a
  missing div term or a missing grad-U term makes the mean drift by many
  standard errors).

Dynamics that hold the centre of mass sample the canonical distribution on
the plane sum_i m_i x_i = const. The local field is a function of
interatomic vectors only, so its Stein sum is translation invariant and has the
same mean on that plane as on the equal-weight plane sum_i x_i = const, where
its projection g_i - mean_j g_j is tangent, has the same divergence (the
derivatives of g along a common translation vanish) and the same g . grad U
(the forces sum to zero): the sum needs no correction. The displacement fields
are not translation invariant; they are tangent to the mass-weighted plane and
lose one unit of divergence per axis there (``global_stein``'s ndiv).

Units: positions and displacements in A, forces in eV/A, energies in eV,
beta in 1/eV.
"""

from __future__ import annotations

import numpy as np

NB = 32


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def min_image(d, box):
    box = np.asarray(box, dtype=float)
    return d - box * np.rint(d / box)


def descriptors(pos, types, box, rc=6.8, mu0=2.0, dmu=0.15, width=0.15, axes=False):
    """D [N, 64], J [3, N, 64] = d D_i / d x_{i,a} and L [N, 64] =
    sum_a d^2 D_i / d x_{i,a}^2, brute force over pairs; with axes also the
    axis terms La [3, N, 64] = d^2 D_i / d x_{i,a}^2.
    phi_k(r) = exp(-(r - mu_k)^2 / (2 w^2)) (cos(pi r / rc) + 1) / 2; the
    Laplacian of a radial function is phi'' + 2 phi' / r, of which axis a
    holds (phi'' - phi'/r)(d_a/r)^2 + phi'/r."""
    pos = np.asarray(pos, dtype=float)
    N = len(pos)
    mu = mu0 + dmu * np.arange(NB)
    D = np.zeros((N, 2 * NB))
    J = np.zeros((3, N, 2 * NB))
    L = np.zeros((N, 2 * NB))
    La = np.zeros((3, N, 2 * NB))
    for i in range(N):
        d = min_image(pos[i] - pos, box)
        r = np.sqrt((d ** 2).sum(1))
        m = (r < rc) & (np.arange(N) != i)
        d, r, t = d[m], r[m], np.asarray(types)[m]
        u = r[:, None] - mu[None, :]
        g = np.exp(-0.5 * u ** 2 / width ** 2)
        c = np.cos(np.pi * r / rc)[:, None]
        sn = np.sin(np.pi * r / rc)[:, None]
        fc = 0.5 * (c + 1.0)
        dfc = -0.5 * np.pi / rc * sn
        d2fc = -0.5 * (np.pi / rc) ** 2 * c
        ui = u / width ** 2
        phi = g * fc
        d1 = g * (dfc - ui * fc)
        d2 = g * ((ui ** 2 - 1.0 / width ** 2) * fc - 2.0 * ui * dfc + d2fc)
        w = d1 / r[:, None]
        lap = d2 + 2.0 * w
        for s in (0, 1):
            sel = t == s
            D[i, s * NB:(s + 1) * NB] = phi[sel].sum(0)
            L[i, s * NB:(s + 1) * NB] = lap[sel].sum(0)
            for a in range(3):
                J[a, i, s * NB:(s + 1) * NB] = (w[sel] * d[sel, a:a + 1]).sum(0)
                cos2 = (d[sel, a] / r[sel]) ** 2
                La[a, i, s * NB:(s + 1) * NB] = ((d2[sel] - w[sel]) * cos2[:, None] + w[sel]).sum(0)
    return (D, J, L, La) if axes else (D, J, L)


def dense_forward(X, T, W, b, act, emulate_bf16=False):
    """Y = act(X W + b), Yt[a] = act'(X W + b) (T[a] W). With emulate_bf16 the
    operands of both products are rounded to bfloat16 as the MFMA path does."""
    if emulate_bf16:
        X, T, W = bf16_round(X), bf16_round(T), bf16_round(W)
    Z = X @ W + b
    Zt = np.einsum("amk,kn->amn", T, W)
    if act == 0:
        Y = np.tanh(Z)
        s = 1.0 - Y ** 2
    else:
        Y, s = Z, np.ones_like(Z)
    return Y, s[None] * Zt


def net_forward(params, D, J, scale=0.1, emulate_bf16=False):
    """The control variate's network: 64 -> 64 -> 64 -> F x 64 (tanh, tanh,
    identity); returns the weights w [M, F, 64] of the F fields and their
    tangents wt [3, M, F, 64] along each atom's own coordinates."""
    W1, b1, W2, b2, W3, b3 = params
    h, ht = dense_forward(D * scale, J * scale, W1, b1, 0, emulate_bf16)
    if emulate_bf16:
        h, ht = h.astype(np.float32), ht.astype(np.float32)
    h, ht = dense_forward(h, ht, W2, b2, 0, emulate_bf16)
    if emulate_bf16:
        h, ht = h.astype(np.float32), ht.astype(np.float32)
    o, ot = dense_forward(h, ht, W3, b3, 1, emulate_bf16)
    M = len(D)
    return o.reshape(M, -1, 2 * NB), ot.reshape(3, M, -1, 2 * NB)


def fields(w, wt, J, L):
    """The fields c_{f,i} = sum_k w_{f,ik} grad_i D_ik (sums of radial vectors
    along the bonds of atom i, equivariant) and their divergences along each
    atom's own coordinates:
      div_i c_{f,i} = sum_k [ w_{f,ik} L_ik + sum_a (d w_{f,ik} / d x_{i,a}) J_{a,ik} ].
    Returns c [M, F, 3] and div [M, F]."""
    c = np.einsum("mfk,amk->mfa", w, J)
    div = np.einsum("mfk,mk->mf", w, L) + np.einsum("amfk,amk->mf", wt, J)
    return c, div


def field_sums(c, div, F, beta, G=None):
    """Per field f of one frame: A_f = sum_i div_i c_f,i + beta c_f,i . F_i (a
    Stein sum), C_f = sum_i c_f,i . F_i and E_f = sum_i c_f,i . G_i."""
    cF = np.einsum("mfa,ma->mf", c, F)
    E = np.einsum("mfa,ma->mf", c, G).sum(0) if G is not None else np.zeros(c.shape[1])
    return (div + beta * cF).sum(0), cF.sum(0), E


def net_features(A, C, E, dh, out_scale=1.0):
    """Per head j (fields 2j, 2j+1): the Stein sums of the plain field and of
    the field times (h - h0), dh A + E (E = sum_i c_i . grad_i h). The sums of
    fields times (U - U0) are dU A - C (grad_i U = -F_i). Returns [3, 2]."""
    A, E = (out_scale * np.asarray(x) for x in (A, E))
    return np.stack([A[0::2], dh * A[1::2] + E[1::2]], axis=1)


def proxy(D, types, V):
    """h = sum_i V[t_i] . D_i: a pair sum sum_{i != j} v_{t_i, t_j}(r_ij) with
    v_{c,t}(r) = sum_m V[c, t*32 + m] phi_m(r)."""
    return float(np.einsum("ik,ik->", D, V[np.asarray(types)]))


def proxy_grad(J, types, V):
    """grad_k h [N, 3]: atom k's own descriptors give V[t_k] . J_k; each
    neighbour j's channel of k's species gives V[t_j, t_k*32 + m] phi_m'(r) r_hat,
    which, summed over the neighbours of one species t, is V[t, t_k*32 + m]
    times atom k's channel t of J."""
    t = np.asarray(types)
    Wg = V.copy()
    for c in range(2):
        for u in range(2):
            Wg[c, u * NB:(u + 1) * NB] += V[u, c * NB:(c + 1) * NB]
    return np.einsum("amk,mk->ma", J, Wg[t])


def h_stein(u, F, h, h0, gh, beta, ndiv):
    """Stein sums of the fields (h - h0) e_a u_ib, [9]:
    (h - h0) S_ab + sum_k (grad_k h)_a u_kb."""
    S = ndiv * np.eye(3) + beta * np.einsum("ia,ib->ab", F, u)
    return ((h - h0) * S + np.einsum("ka,kb->ab", gh, u)).ravel()


def global_stein(u, F, U, U0, beta, ndiv=None):
    """Zero-mean Stein sums of the displacement fields, per frame.

    u: displacements from the lattice sites [N, 3]; F: forces [N, 3]; U: the
    configurational energy; U0: any constant; ndiv: the divergence of the
    field e_a u_ia on the configuration space, N when every atom moves freely
    and N - 1 when the centre of mass is held (the fields are tangent to that
    constraint because sum_i m_i u_i = 0, and the divergence along the
    constraint drops the one normal direction per axis). Fields and sums:
      e_a u_ib          S_ab = ndiv d_ab + beta sum_i u_ib F_ia
      (U - U0) e_a u_ib T_ab = (U - U0) S_ab - sum_i F_ia u_ib
      (U - U0)^2 e_a u_ib Q_ab = (U - U0)^2 S_ab - 2 (U - U0) sum_i F_ia u_ib
    (div of h(U) v = h'(U) grad U . v + h(U) div v, grad_i U = -F_i).
    Returns [27]: S, T, Q flattened row-major over (a, b)."""
    N = len(u)
    nd = N if ndiv is None else ndiv
    M = np.einsum("ia,ib->ab", F, u)          # M[a, b] = sum_i F_ia u_ib
    S = nd * np.eye(3) + beta * M
    dU = U - U0
    T = dU * S - M
    Q = dU ** 2 * S - 2.0 * dU * M
    return np.concatenate([S.ravel(), T.ravel(), Q.ravel()])


# ---------------------------------------------------------------------------
# stress proxy and its zero-variance fields
# ---------------------------------------------------------------------------
VOIGT = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))


def _ordered_pairs(pos, box, rc):
    """Ordered pairs (i, j), i != j, r_ij < rc: i, j, d = x_i - x_j (minimum
    image), r."""
    pos = np.asarray(pos, dtype=float)
    d = pos[:, None, :] - pos[None, :, :]
    d = min_image(d, box)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    i, j = np.nonzero(r2 < rc * rc)
    dv = d[i, j]
    return i, j, dv, np.sqrt(r2[i, j])


def _gauss(r, rc, mu0, dmu, width):
    """g_k(r) [P, 32], fc, fc', fc''."""
    mu = mu0 + dmu * np.arange(NB)
    g = np.exp(-(r[:, None] - mu) ** 2 / (2.0 * width ** 2))
    a = np.pi / rc
    return g, 0.5 * (np.cos(a * r) + 1.0), -0.5 * a * np.sin(a * r), -0.5 * a * a * np.cos(a * r)


def stress_basis_sums(pos, types, box, rc=6.8, mu0=2.0, dmu=0.15, width=0.15):
    """Z [3, 32, 6]: over ordered pairs of species pair p = t_i + t_j,
    sum fc(r) g_k(r) n_a n_b, Voigt order xx, yy, zz, yz, xz, xy."""
    t = np.asarray(types)
    i, j, d, r = _ordered_pairs(pos, box, rc)
    g, fc, _, _ = _gauss(r, rc, mu0, dmu, width)
    n = d / r[:, None]
    Z = np.zeros((3, NB, 6))
    p = t[i] + t[j]
    for v, (a, b) in enumerate(VOIGT):
        w = g * (fc * n[:, a] * n[:, b])[:, None]
        for q in range(3):
            Z[q, :, v] = w[p == q].sum(0)
    return Z


def stress_proxy(pos, types, box, coef, rc=6.8, mu0=2.0, dmu=0.15, width=0.15):
    """Per atom: the share of h_v (sum_j psi n_a n_b) [N, 6], the Laplacian of
    h_v along the atom's coordinates [N, 6] and grad h_v [N, 6, 3], for
    h_v = sum over ordered pairs of psi_p(r) n_a n_b, psi_p = fc sum_k coef[p, k] g_k."""
    pos = np.asarray(pos, dtype=float)
    t = np.asarray(types)
    N = len(pos)
    i, j, d, r = _ordered_pairs(pos, box, rc)
    g, fc, dfc, ddfc = _gauss(r, rc, mu0, dmu, width)
    mu = mu0 + dmu * np.arange(NB)
    q = (r[:, None] - mu) / width ** 2
    c = np.asarray(coef, dtype=float)[t[i] + t[j]]
    s0 = (c * g).sum(1)
    s1 = -(c * g * q).sum(1)
    s2 = (c * g * (q * q - 1.0 / width ** 2)).sum(1)
    psi = fc * s0
    dpsi = dfc * s0 + fc * s1
    ddpsi = ddfc * s0 + 2.0 * dfc * s1 + fc * s2
    n = d / r[:, None]
    hp = np.zeros((N, 6))
    lap = np.zeros((N, 6))
    grad = np.zeros((N, 6, 3))
    eye = np.eye(3)
    for v, (a, b) in enumerate(VOIGT):
        nab = n[:, a] * n[:, b]
        np.add.at(hp[:, v], i, psi * nab)
        lp = (ddpsi + 2.0 * dpsi / r - 6.0 * psi / r ** 2) * nab + (2.0 * psi / r ** 2) * eye[a, b]
        # both atoms of an ordered pair see the Laplacian in d
        np.add.at(lap[:, v], i, lp)
        np.add.at(lap[:, v], j, lp)
        T = (dpsi * nab)[:, None] * n + (psi / r)[:, None] * (
            eye[a][None, :] * n[:, b:b + 1] + eye[b][None, :] * n[:, a:a + 1]
            - 2.0 * nab[:, None] * n)
        # d/dx_i of the pair term is T(d), d/dx_j is -T(d); T is odd in d
        np.add.at(grad[:, v, :], i, T)
        np.add.at(grad[:, v, :], j, -T)
    return hp, lap, grad


def zv_stein(h, h0, lap, grad, F, beta):
    """Stein sums of the zero-variance fields of the stress proxy h [6] (h0 any
    constant [6]; lap [6] = sum_m lap_m h, grad [N, 6, 3], F [N, 3] in eV/A):
      grad h_a:          A_a  = lap h_a + beta sum_m grad_m h_a . F_m
      (h_a - h0_a) grad h_b:  B_ab = sum_m grad_m h_a . grad_m h_b + (h_a - h0_a) A_b
    Returns [42]: A (6), then B row-major (36)."""
    A = np.asarray(lap) + beta * np.einsum("mac,mc->a", grad, F)
    GG = np.einsum("mac,mbc->ab", grad, grad)
    B = GG + (np.asarray(h) - np.asarray(h0))[:, None] * A[None, :]
    return np.concatenate([A, B.ravel()])
