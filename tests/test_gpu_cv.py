"""Control-variate kernels (include/mdelast_cv.h) against the CPU restatement
(oracle/cv_oracle.py) and an FP64 autograd reference, and the trained variate
on molecular-dynamics frames of the device loop.

Tolerances: the products round their operands to bfloat16 and accumulate in
FP32; the oracle rounds the same operands the same way, so forward values and
tangents agree to FP32 accumulation (norm-wise 1e-4 and 2e-4; an operand that
lands on the other side of a bfloat16 rounding boundary moves one element by
2^-9 relative). Gradients are compared with an FP64 reference that does not round,
so they agree to the bfloat16 operand precision (norm-wise 2e-2).
"""

from __future__ import annotations

import numpy as np
import pytest

import cv_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from mdelast_amd import cv  # noqa: E402

DEV = "cuda:0"


def _cell(seed=0, edge=4, amp=0.12, a=3.667):
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    idx = np.indices((edge,) * 3).reshape(3, -1).T
    sites = ((idx[:, None, :] + basis[None]) * a).reshape(-1, 3)
    rs = np.random.RandomState(seed)
    pos = sites + rs.normal(scale=amp, size=sites.shape)
    types = (rs.rand(len(sites)) < 0.25).astype(np.int32)
    return sites, pos, types, np.full(3, edge * a)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_descriptors_match_the_oracle():
    _, pos, types, box = _cell()
    print("descriptors: 256 atoms", flush=True)
    D, J, L = cv.descriptors(torch.tensor(pos, device=DEV), torch.tensor(types, device=DEV), box)
    Do, Jo, Lo = O.descriptors(pos, types, box)
    assert np.allclose(D.cpu().numpy(), Do, rtol=2e-5, atol=2e-6 * np.abs(Do).max())
    assert np.allclose(J.cpu().numpy(), Jo, rtol=2e-5, atol=2e-6 * np.abs(Jo).max())
    assert np.allclose(L.cpu().numpy(), Lo, rtol=2e-5, atol=2e-6 * np.abs(Lo).max())


def test_driver_descriptors_match_the_oracle():
    """The list-based descriptors of a 500-atom brass cell after some steps,
    against the brute-force oracle on the driver's own positions."""
    print("driver descriptors: brass 500 atoms", flush=True)
    drv, sites, types = cv.brass_driver(edge=5)
    drv.run(300, sample_stride=300)
    La = torch.empty((3, drv.N, 64), device=DEV)
    D, J, L, P2, F2 = cv.driver_descriptors(drv, La=La)
    pos = drv.positions_array() / cv.U.angstrom
    box = np.diag(drv.box.h) / cv.U.angstrom
    Do, Jo, Lo, Lao = O.descriptors(pos, types, box, axes=True)
    assert np.allclose(La.cpu().numpy(), Lao, rtol=2e-5, atol=2e-6 * np.abs(Lao).max())
    assert np.allclose(D.cpu().numpy(), Do, rtol=2e-5, atol=2e-6 * np.abs(Do).max())
    assert np.allclose(J.cpu().numpy(), Jo, rtol=2e-5, atol=2e-6 * np.abs(Jo).max())
    assert np.allclose(L.cpu().numpy(), Lo, rtol=2e-5, atol=2e-6 * np.abs(Lo).max())
    F = cv.driver_atoms(drv, "forces").cpu().numpy()
    assert np.array_equal(F, drv.forces_array())
    P = cv.driver_atoms(drv, "positions").cpu().numpy()
    assert np.array_equal(P, drv.positions_array())
    assert np.array_equal(P2.cpu().numpy(), P) and np.array_equal(F2.cpu().numpy(), F)


def test_descriptors_refuse_a_cutoff_beyond_half_the_box():
    _, pos, types, box = _cell(edge=3)
    with pytest.raises(Exception, match="half the box"):
        cv.descriptors(torch.tensor(pos, device=DEV), torch.tensor(types, device=DEV), box)


@pytest.mark.parametrize("R,K,N,transB", [(100, 64, 48, 0), (37, 64, 64, 1), (16, 16, 16, 0),
                                          (2049, 64, 16, 1), (300, 384, 64, 1),
                                          (70000, 64, 384, 0), (40, 48, 32, 0)])
def test_gemm_matches_bf16_operands_in_fp64(R, K, N, transB):
    rs = np.random.RandomState(R)
    A = rs.normal(size=(R, K)).astype(np.float32)
    W = rs.normal(size=(N, K) if transB else (K, N)).astype(np.float32)
    b = rs.normal(size=N).astype(np.float32)
    print(f"gemm {R}x{K}x{N} transB={transB}", flush=True)
    C = cv._gemm(torch.tensor(A, device=DEV), torch.tensor(W, device=DEV),
                 torch.tensor(b, device=DEV), transB=bool(transB)).cpu().numpy()
    Wr = O.bf16_round(W)
    ref = O.bf16_round(A) @ (Wr.T if transB else Wr) + b
    scale = np.abs(O.bf16_round(A)) @ np.abs(Wr.T if transB else Wr)
    assert np.all(np.abs(C - ref) <= 1e-6 * scale + 1e-6)


@pytest.mark.parametrize("R,rows", [(1000, 512), (48, 16), (33, 32)])
def test_gemm_tn(R, rows):
    rs = np.random.RandomState(R)
    A = rs.normal(size=(R, 64)).astype(np.float32)
    B = rs.normal(size=(R, 32)).astype(np.float32)
    print(f"gemm_tn R={R} rows={rows}", flush=True)
    C = cv._gemm_tn(torch.tensor(A, device=DEV), torch.tensor(B, device=DEV), rows=rows)
    ref = O.bf16_round(A).T @ O.bf16_round(B)
    scale = np.abs(O.bf16_round(A)).T @ np.abs(O.bf16_round(B))
    assert np.all(np.abs(C.cpu().numpy() - ref) <= 1e-6 * scale + 1e-6)


@pytest.mark.parametrize("R,N", [(5000, 64), (33, 384), (1024, 16), (1, 48)])
def test_colsum(R, N):
    B = np.random.RandomState(R).normal(size=(R, N)).astype(np.float32)
    print(f"colsum R={R} N={N}", flush=True)
    out = cv._colsum(torch.tensor(B, device=DEV)).cpu().numpy()
    assert np.allclose(out, B.astype(np.float64).sum(0), rtol=1e-5,
                       atol=1e-6 * np.abs(B).sum(0).max())


def _net(seed=3):
    """A network with weights of order one; every draw seeded here (not from
    the device generator, whose state depends on what ran before)."""
    net = cv.SteinNet(seed=seed, device=DEV)
    rs = np.random.RandomState(100 + seed)
    with torch.no_grad():
        net.W3.mul_(30.0)
        net.b1.copy_(torch.tensor(rs.normal(0, 0.1, net.b1.shape), dtype=torch.float32))
        net.b3.copy_(torch.tensor(rs.normal(0, 0.1, net.b3.shape), dtype=torch.float32))
    return net


def test_network_and_field_sums_match_the_oracle():
    _, pos, types, box = _cell(seed=1)
    D, J, L = O.descriptors(pos, types, box)
    net = _net()
    t = [torch.tensor(x, dtype=torch.float32, device=DEV) for x in (D, J, L)]
    print("network forward: 256 atoms", flush=True)
    w, wt = net.forward(t[0], t[1])
    wo, wto = O.net_forward(net.numpy_params(), D.astype(np.float32), J.astype(np.float32),
                            emulate_bf16=True)
    assert _rel(w.detach().cpu().numpy(), wo) < 1e-4
    # the tangents pass their own operand (T W at each layer) through the
    # bfloat16 rounding as well, so twice as many operands can land on the
    # other side of a rounding boundary
    assert _rel(wt.detach().cpu().numpy(), wto) < 2e-4
    F = np.random.RandomState(2).normal(size=(len(pos), 3))
    G = np.random.RandomState(3).normal(size=(len(pos), 3))
    A, Cs, E = net.field_sums(t[0], t[1], t[2],
                              torch.tensor(F, dtype=torch.float32, device=DEV),
                              torch.tensor(G, dtype=torch.float32, device=DEV), 38.7, 1)
    c, div = O.fields(wo, wto, J.astype(np.float32), L.astype(np.float32))
    Ao, Co, Eo = O.field_sums(c, div, F, 38.7, G)
    assert _rel(A.detach().cpu().numpy()[0], net.out_scale * Ao) < 1e-3
    assert _rel(Cs.detach().cpu().numpy()[0], net.out_scale * Co) < 1e-3
    assert _rel(E.detach().cpu().numpy()[0], net.out_scale * Eo) < 1e-3
    one = torch.tensor([0.7], dtype=torch.float64, device=DEV)
    nx = cv.net_features(A, Cs, E, -one)
    ref = O.net_features(Ao, Co, Eo, -0.7, net.out_scale)
    assert np.allclose(nx.detach().cpu().numpy()[0], ref, rtol=1e-3,
                       atol=1e-3 * np.abs(ref).max())


def _torch_reference(params, D, J, L, F, G, beta, out_scale, scale=0.1):
    """FP64 autograd reference of the field sums (CPU, no rounding)."""
    W1, b1, W2, b2, W3, b3 = params
    X, T = D * scale, J * scale

    def layer(X, T, W, b, act):
        Z = X @ W + b
        Zt = torch.stack([T[a] @ W for a in range(3)])
        if act == 0:
            Y = torch.tanh(Z)
            return Y, (1 - Y ** 2)[None] * Zt
        return Z, Zt

    h, ht = layer(X, T, W1, b1, 0)
    h, ht = layer(h, ht, W2, b2, 0)
    o, ot = layer(h, ht, W3, b3, 1)
    M = D.shape[0]
    w, wt = o.reshape(M, 6, 64), ot.reshape(3, M, 6, 64)
    c = (w[None] * J[:, :, None, :]).sum(3)                      # [3, M, 6]
    div = (w * L[:, None, :]).sum(2) + (wt * J[:, :, None, :]).sum((0, 3))
    cF = (c * F.T[:, :, None]).sum(0)
    cG = (c * G.T[:, :, None]).sum(0)
    return (out_scale * (div + beta * cF).sum(0), out_scale * cF.sum(0),
            out_scale * cG.sum(0))


def test_backward_matches_fp64_autograd():
    _, pos, types, box = _cell(seed=4)
    D, J, L = O.descriptors(pos, types, box)
    F = np.random.RandomState(5).normal(size=(len(pos), 3))
    G = np.random.RandomState(8).normal(size=(len(pos), 3))
    net = _net(seed=6)
    wts = torch.tensor(np.random.RandomState(7).normal(size=18))
    print("network backward: 256 atoms", flush=True)
    t = [torch.tensor(x, dtype=torch.float32, device=DEV) for x in (D, J, L)]
    A, Cs, E = net.field_sums(t[0], t[1], t[2],
                              torch.tensor(F, dtype=torch.float32, device=DEV),
                              torch.tensor(G, dtype=torch.float32, device=DEV), 38.7, 1)
    out = torch.cat([A[0], Cs[0], E[0]]).double().cpu()
    loss = (out * wts).sum() + (out ** 2).sum() * 1e-3
    loss.backward()
    grads = [p.grad.detach().double().cpu().numpy() for p in net.parameters()]
    ref_p = [torch.tensor(O.bf16_round(p) if p.ndim == 2 else p, requires_grad=True)
             for p in net.numpy_params()]
    Ar, Cr, Er = _torch_reference(ref_p, torch.tensor(D), torch.tensor(J), torch.tensor(L),
                                  torch.tensor(F), torch.tensor(G), 38.7, net.out_scale)
    outr = torch.cat([Ar, Cr, Er])
    lr = (outr * wts).sum() + (outr ** 2).sum() * 1e-3
    lr.backward()
    assert _rel(out.detach().numpy(), outr.detach().numpy()) < 2e-2
    for k, (g, p) in enumerate(zip(grads, ref_p)):
        assert _rel(g, p.grad.numpy()) < 2e-2, k


def test_stein_identity_and_fallback_on_md_frames():
    """500-atom alpha-brass NVT: the trained control variate keeps a zero mean
    on a fresh trajectory (Stein); the batched in-loop evaluation from the
    driver's device state gives the stored path's controls; switched-off heads
    reproduce the plain estimate exactly."""
    print("brass 500 atoms: frames", flush=True)
    drv, sites, types = cv.brass_driver(edge=5)
    drv.run(4000, sample_stride=4000)
    fr = cv.collect_frames(drv, 500, 20, 300.0, sites=sites)
    train, valid = cv.subset(fr, slice(0, 400)), cv.subset(fr, slice(400, None))
    cvm = cv.ControlVariate(sites, fr.box, 300.0, types, device=DEV)
    print("brass 500 atoms: training", flush=True)
    info = cvm.fit(train, valid, epochs=2)
    assert info["proxy_r2"] > 0.5
    d, _, _ = cv.brass_driver(edge=5, vel_seed=999)
    d.run(4000, sample_stride=4000)
    ev = cv.collect_frames(d, 1280, 20, 300.0, sites=sites)
    s_all = cvm.controls(ev, enabled_only=False)
    se = np.array([cv.block_bootstrap_se(s_all[:, h]) for h in range(3)])
    print("mean s", s_all.mean(0), "se", se, flush=True)
    assert np.all(np.abs(s_all.mean(0)) < 4.5 * se)
    # batched in-loop evaluation == stored frames
    batch = cv.ControlBatch(cvm, d.N, B=4)
    ref = cv.collect_frames(d, 0, 20, 300.0)
    eps = []
    st = {q: [] for q in ("D", "J", "L", "pos", "F")}
    for k in range(6):
        rec = d.run(20, sample_stride=20)[-1]
        eps.append(cv._record_scalars(rec, d.N)[0])
        batch.push(d, eps[-1])
        for q, v in zip(("D", "J", "L", "pos", "F"), cv.frame_inputs(d)):
            st[q].append(v)
    s_loop = batch.result()
    ref.store = {q: torch.stack(v) for q, v in st.items()}
    ref.epot = np.asarray(eps)
    s_ref = cvm.controls(ref, enabled_only=False)
    assert s_loop.shape == (6, 3)
    # the same inputs; FP32 sums over the atoms of a frame may round in
    # another order for another batch shape
    print("in-loop", s_loop.ravel(), "stored", s_ref.ravel(), flush=True)
    assert _rel(s_loop, s_ref) < 1e-5
    # inputs captured inside the runs (ControlBatch.run) == stored frames of an
    # identical trajectory; the capture leaves the trajectory untouched
    d1, _, _ = cv.brass_driver(edge=5, vel_seed=777)
    d2, _, _ = cv.brass_driver(edge=5, vel_seed=777)
    d1.run(200, sample_stride=200)
    d2.run(200, sample_stride=200)
    cap = cv.ControlBatch(cvm, d1.N, B=4)
    recs = cap.run(d1, 10, 20)
    s_cap = cap.result()
    st = {q: [] for q in ("D", "J", "L", "pos", "F")}
    recs2 = []
    for k in range(10):
        recs2.append(d2.run(20, sample_stride=20)[-1])
        for q, v in zip(("D", "J", "L", "pos", "F"), cv.frame_inputs(d2)):
            st[q].append(v)
    assert np.array_equal(recs, np.array(recs2))
    ref2 = cv.collect_frames(d2, 0, 20, 300.0)
    ref2.store = {q: torch.stack(v) for q, v in st.items()}
    ref2.epot = np.array([cv._record_scalars(r, d2.N)[0] for r in recs2])
    s_ref2 = cvm.controls(ref2, enabled_only=False)
    print("captured", s_cap.ravel(), "stored", s_ref2.ravel(), flush=True)
    assert s_cap.shape == (10, 3)
    assert _rel(s_cap, s_ref2) < 1e-5
    cvm.enabled[:] = False
    Cp, sep = cv.estimate(ev)
    Cc, sec = cv.estimate(ev, cvm.controls(ev))
    assert np.array_equal(Cp, Cc) and np.array_equal(sep, sec)


def test_in_loop_controls_with_stress_fields_use_the_captured_pairs():
    """A variate with the stress proxy, in the loop of a run (ControlBatch.run:
    inputs and pair lists captured inside the run) against the same variate on
    the stored frames of an identical trajectory (all-pairs stress kernels)."""
    print("brass 500 atoms: stress variate in the loop", flush=True)
    drv, sites, types = cv.brass_driver(edge=5)
    drv.run(2000, sample_stride=2000)
    fr = cv.collect_frames(drv, 300, 20, 300.0, sites=sites)
    cvm = cv.ControlVariate(sites, fr.box, 300.0, types, device=DEV, use_net=False,
                            use_stress=True)
    cvm.fit(cv.subset(fr, slice(0, 240)), cv.subset(fr, slice(240, None)))
    d1, _, _ = cv.brass_driver(edge=5, vel_seed=555)
    d2, _, _ = cv.brass_driver(edge=5, vel_seed=555)
    d1.run(200, sample_stride=200)
    d2.run(200, sample_stride=200)
    cap = cv.ControlBatch(cvm, d1.N, B=4)
    assert cap.pairs is not None
    recs = cap.run(d1, 10, 20)
    s_cap = cap.result()
    st = {q: [] for q in ("D", "J", "L", "pos", "F")}
    recs2 = []
    for k in range(10):
        recs2.append(d2.run(20, sample_stride=20)[-1])
        for q, v in zip(("D", "J", "L", "pos", "F"), cv.frame_inputs(d2)):
            st[q].append(v)
    assert np.array_equal(recs, np.array(recs2))
    ref = cv.collect_frames(d2, 0, 20, 300.0)
    ref.store = {q: torch.stack(v) for q, v in st.items()}
    ref.epot = np.array([cv._record_scalars(r, d2.N)[0] for r in recs2])
    s_ref = cvm.controls(ref, enabled_only=False)
    print("captured", s_cap.ravel(), "stored", s_ref.ravel(), flush=True)
    assert s_cap.shape == (10, 3)
    assert _rel(s_cap, s_ref) < 1e-5


def test_learned_fields_are_kept_only_where_they_help():
    """The fit measures, per head and on held-out frames, the plain variance,
    the variance with the global fields and the variance with the learned
    fields added. A head keeps its learned fields only where they cut its
    variance by 5 %, and the network only if that cuts the worst head's
    controlled/plain ratio by 5 %: the variate never does worse on the
    held-out frames than the global fields alone or than plain."""
    print("brass 500 atoms: network gain", flush=True)
    drv, sites, types = cv.brass_driver(edge=5, vel_seed=17)
    drv.run(4000, sample_stride=4000)
    fr = cv.collect_frames(drv, 600, 20, 300.0, sites=sites)
    train, valid = cv.subset(fr, slice(0, 480)), cv.subset(fr, slice(480, None))
    cvm = cv.ControlVariate(sites, fr.box, 300.0, types, device=DEV, seed=3)
    info = cvm.fit(train, valid, epochs=4)
    print({k: v for k, v in info.items() if k != "history"}, flush=True)
    vp, vg, vf, vc = (np.asarray(info[k]) for k in ("var_plain_valid", "var_global_valid",
                                                    "var_fields_valid", "var_cv_valid"))
    used = np.asarray(info["uses_learned_fields"])
    on = np.asarray(info["enabled"])
    assert np.all(vc <= vg)
    assert np.array_equal(on, vc < vp)
    assert np.array_equal(vc[on], np.where(used, vf, vg)[on])
    assert np.all(vf[used] < 0.95 * vg[used])
    if used.any():
        def worst(v):
            return np.max(np.where(v < vp, v / vp, 1.0))
        assert worst(vc) < 0.95 * worst(vg)
    assert cvm.use_net == bool(used.any())
    # the kept controls reproduce the held-out variance the fit reported
    s = cvm.controls(valid, enabled_only=False)
    assert np.allclose((valid.y() - s).var(0), vc, rtol=1e-6)
    assert len(info["history"]) == 4


def test_stress_kernels_match_the_oracle():
    """The stress proxy's basis sums, pair sums, Laplacians and per-atom
    gradients on three frames at once against the oracle (FP64 both)."""
    frames = [_cell(seed=s) for s in (21, 22, 23)]
    types = frames[0][2]
    box = frames[0][3]
    pos = np.stack([f[1] for f in frames])
    coef = np.random.RandomState(4).normal(size=(3, 32)) * 0.3
    t = torch.tensor(pos, dtype=torch.float64, device=DEV)
    ty = torch.tensor(types, dtype=torch.int32, device=DEV)
    print("stress proxy kernels: 3 x 256 atoms", flush=True)
    Z = cv.stress_basis_sums(t, ty, box).cpu().numpy()
    hp, lap, grad = cv.stress_proxy(t, ty, box, cv.Basis(),
                                    torch.tensor(coef, dtype=torch.float64, device=DEV))
    hp, lap, grad = (x.cpu().numpy() for x in (hp, lap, grad))
    for b in range(3):
        Zo = O.stress_basis_sums(pos[b], types, box)
        hpo, lapo, grado = O.stress_proxy(pos[b], types, box, coef)
        assert _rel(Z[b], Zo) < 1e-12
        assert _rel(hp[b], hpo.sum(0)) < 1e-12
        assert _rel(lap[b], lapo.sum(0)) < 1e-12
        assert _rel(grad[b], grado) < 1e-12
    # the per-frame field sums of one pass (mdl_cv_stress_fields) against the
    # oracle's per-atom terms contracted on the host
    Fr = np.random.RandomState(5).normal(size=pos.shape)
    sf = cv.stress_field_sums(t, torch.tensor(Fr, dtype=torch.float64, device=DEV), ty, box,
                              cv.Basis(), torch.tensor(coef, dtype=torch.float64, device=DEV))
    sf = sf.cpu().numpy()
    for b in range(3):
        hpo, lapo, grado = O.stress_proxy(pos[b], types, box, coef)
        gF = np.einsum("mvc,mc->v", grado, Fr[b])
        gg = np.einsum("mac,mbc->ab", grado, grado)
        assert _rel(sf[b, :6], hpo.sum(0)) < 1e-12
        assert _rel(sf[b, 6:12], lapo.sum(0)) < 1e-12
        assert _rel(sf[b, 12:18], gF) < 1e-12
        assert _rel(sf[b, 18:], np.array([gg[a, c] for a, c in cv.SF_PAIRS])) < 1e-12


def test_captured_pair_lists_give_the_all_pairs_sums():
    """The stress kernels on the pair lists captured in the driver's run (its
    compacted pairs within r_c, list order) against the same kernels searching
    all atoms: the same pairs, another summation order."""
    print("brass 500 atoms: 24 frames with pair lists", flush=True)
    drv, sites, types = cv.brass_driver(edge=5)
    drv.run(2000, sample_stride=2000)
    fr = cv.capture_frames(drv, 24, 20, 300.0)
    cnt = fr.store["ncnt"].cpu().numpy()
    assert cnt.min() > 50 and cnt.max() <= cv.PAIR_WIDTH
    posA = fr.store["pos"] / cv.U.angstrom
    F = fr.store["F"] * (cv.U.angstrom / cv.U.ev)
    ty = torch.tensor(types, dtype=torch.int32, device=DEV)
    pairs = (fr.store["nbr"], fr.store["ncnt"])
    Z0 = cv.stress_basis_sums(posA, ty, fr.box).cpu().numpy()
    Z1 = cv.stress_basis_sums(posA, ty, fr.box, pairs=pairs).cpu().numpy()
    assert _rel(Z1, Z0) < 1e-12
    coef = torch.tensor(np.random.RandomState(6).normal(size=(3, 32)) * 0.3, dtype=torch.float64,
                        device=DEV)
    s0 = cv.stress_field_sums(posA, F, ty, fr.box, cv.Basis(), coef).cpu().numpy()
    s1 = cv.stress_field_sums(posA, F, ty, fr.box, cv.Basis(), coef, pairs).cpu().numpy()
    for a, b in ((0, 6), (6, 12), (12, 18), (18, 39)):
        assert _rel(s1[:, a:b], s0[:, a:b]) < 1e-12


def test_stress_fields_on_md_frames():
    """500-atom alpha-brass NVT, frames captured inside one run: the stress
    proxy fits the Born stress, the variate's 42 stress-field sums equal the
    oracle's on stored frames, and the fitted variate keeps a zero mean on a
    fresh trajectory (Stein)."""
    print("brass 500 atoms: captured frames", flush=True)
    drv, sites, types = cv.brass_driver(edge=5)
    drv.run(4000, sample_stride=4000)
    fr = cv.capture_frames(drv, 800, 20, 300.0)
    train, valid = cv.subset(fr, slice(0, 640)), cv.subset(fr, slice(640, None))
    cvm = cv.ControlVariate(sites, fr.box, 300.0, types, device=DEV, use_net=False,
                            use_stress=True)
    info = cvm.fit(train, valid, epochs=0)
    r2 = cvm.fit_stress_proxy(train)
    print("stress proxy R^2", r2, "held-out ratios",
          np.asarray(info["var_cv_valid"]) / np.asarray(info["var_plain_valid"]), flush=True)
    assert np.all(r2 > 0.9)
    coef = cvm.sp_coef.cpu().numpy()
    for k in (3, 500):
        posA = fr.store["pos"][k].cpu().numpy() / cv.U.angstrom
        F = fr.store["F"][k].cpu().numpy() * (cv.U.angstrom / cv.U.ev)
        hp, lap, grad = O.stress_proxy(posA, types, fr.box, coef)
        ref = O.zv_stein(hp.sum(0) + cvm.sp_off, cvm.sp_h0, lap.sum(0), grad, F, cvm.beta)
        got = cvm.stress_features(fr.store["pos"][k:k + 1],
                                  fr.store["F"][k:k + 1] * (cv.U.angstrom / cv.U.ev))
        assert _rel(got.cpu().numpy()[0], ref) < 1e-10
    print("brass 500 atoms: fresh trajectory", flush=True)
    d, _, _ = cv.brass_driver(edge=5, vel_seed=999)
    d.run(4000, sample_stride=4000)
    ev = cv.capture_frames(d, 1280, 20, 300.0)
    s_all = cvm.controls(ev, enabled_only=False)
    se = np.array([cv.block_bootstrap_se(s_all[:, h], 32) for h in range(3)])
    print("mean s", s_all.mean(0), "se", se, flush=True)
    assert np.all(np.abs(s_all.mean(0)) < 4.5 * se)


def test_crossfit_at_the_4000_atom_configuration():
    """The control variate's own configuration: 4 000-atom alpha-brass (10^3
    cells, 240 Zn), NVT 300 K, frames every 20 steps, captured inside one run
    with their pair lists. On two of its frames the stress kernels (captured
    lists) match the oracle's all-pairs sums; cross-fitted on 600 frames, the
    controls keep a zero mean within their block-bootstrap error, and the
    controlled per-frame variance of C11 (the head that sets the cost) is well
    below the plain one."""
    print("brass 4000 atoms: 600 captured frames", flush=True)
    drv, sites, types = cv.brass_driver(edge=10)
    assert drv.N == 4000
    drv.run(4000, sample_stride=4000)
    fr = cv.capture_frames(drv, 600, 20, 300.0)
    ty = torch.tensor(types, dtype=torch.int32, device=DEV)
    pairs = (fr.store["nbr"], fr.store["ncnt"])
    coef = np.random.RandomState(8).normal(size=(3, 32)) * 0.3
    for k in (0, 599):
        posA = fr.store["pos"][k:k + 1] / cv.U.angstrom
        F = fr.store["F"][k:k + 1] * (cv.U.angstrom / cv.U.ev)
        sl = (pairs[0][k:k + 1], pairs[1][k:k + 1])
        Z = cv.stress_basis_sums(posA, ty, fr.box, pairs=sl).cpu().numpy()[0]
        sf = cv.stress_field_sums(posA, F, ty, fr.box, cv.Basis(),
                                  torch.tensor(coef, dtype=torch.float64, device=DEV),
                                  sl).cpu().numpy()[0]
        p = posA[0].cpu().numpy()
        Fh = F[0].cpu().numpy()
        assert _rel(Z, O.stress_basis_sums(p, types, fr.box)) < 1e-12
        hpo, lapo, grado = O.stress_proxy(p, types, fr.box, coef)
        gg = np.einsum("mac,mbc->ab", grado, grado)
        assert _rel(sf[:6], hpo.sum(0)) < 1e-12
        assert _rel(sf[6:12], lapo.sum(0)) < 1e-12
        assert _rel(sf[12:18], np.einsum("mvc,mc->v", grado, Fh)) < 1e-12
        assert _rel(sf[18:], np.array([gg[a, c] for a, c in cv.SF_PAIRS])) < 1e-12
    print("brass 4000 atoms: cross-fit", flush=True)
    s, fits = cv.crossfit_controls(fr, sites, types, 300.0)
    assert s.shape == (600, 3)
    y = fr.y()
    se = np.array([cv.block_bootstrap_se(s[:, h], 16) for h in range(3)])
    ratio = (y - s).var(0) / y.var(0)
    print("mean s", s.mean(0), "se", se, "variance ratio", ratio, flush=True)
    # a head switched off in both folds has s = 0 exactly
    assert np.all(np.abs(s.mean(0)) <= 4.5 * se)
    assert ratio[0] < 0.4
