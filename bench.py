#!/usr/bin/env python3
"""Throughput of the MI355X EAM step on the large single-instance workload.

Workload (one GPU): three-element Ni-Al-Co fcc random solution, Ni 75 / Al 20 /
Co 5 at.% placed by a seeded permutation, 30x30x30 cells = 108,000 atoms,
a = 3.56 A, Mishin_updated-Ni-Al-Co-2013 EAM tables (quintic), NVT at 600 K,
dt = 1 fs, Nose-Hoover chain of 10, virial and the 21-component Born tensor
accumulated every step, neighbour list r_skin 1.2 rebuilt every 20 steps
(plus whenever an atom pair could have crossed into the cut-off).

A step is one full MD step (md_driver.cpp:563-601 of the reference): NHC,
kick, drift, list check/build, density pass, force pass, kick, NHC and
thermodynamics. value = atoms * steps / wall time of the timed steps, inputs
resident in HBM. Default arithmetic: mixed (FP32 tables and pair algebra,
FP64 positions, chains and accumulators); --precision fp64 runs the all-FP64
path. The line also carries the force kernel's roofline position (algorithmic
bytes / measured kernel time, HIP events on the driver's stream; HBM traffic
from the committed rocprofv3 --pmc summary) and the CPU oracle timed on a
bounded sample of the same workload on one core.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--precision mixed|fp64]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mdelast_amd import md_driver_module as mdm  # noqa: E402
from mdelast_amd import setfl, units as U  # noqa: E402

POTENTIAL = "Mishin_updated-Ni-Al-Co-2013.eam.alloy"
HBM_PEAK_GBS = 8000.0


def build_workload(nside=30, a=3.56, fractions=(0.75, 0.20, 0.05), seed=42, T=600.0, M=10,
                   tau_T=5e-14, vel_seed=0):
    """Positions (CGS), types, box, and MolecularDynamics-style velocities and chain."""
    data = setfl.read_setfl(POTENTIAL)
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    idx = np.indices((nside, nside, nside)).reshape(3, -1).T
    pos = ((idx[:, None, :] + 0.25 + basis[None, :, :]) * a).reshape(-1, 3) - nside * a / 2.0
    N = len(pos)
    counts = [int(round(f * N)) for f in fractions]
    counts[0] = N - sum(counts[1:])
    perm = np.random.RandomState(seed).permutation(N)
    types = np.zeros(N, dtype=np.uint32)
    k = 0
    for t, c in enumerate(counts):
        types[perm[k:k + c]] = t
        k += c
    mass_t = np.array([e.atomic_weight / U.Navogadro for e in data.elements])
    m = mass_t[types]
    rs = np.random.RandomState(vel_seed)
    sig = np.sqrt(U.k_boltz * T / m)
    v = np.stack([rs.normal(0.0, sig), rs.normal(0.0, sig), rs.normal(0.0, sig)], axis=1)
    v -= (v * m[:, None]).sum(0) / m.sum()
    v *= np.sqrt((3.0 * N - 3.0) * U.k_boltz * T / 2.0 / (0.5 * np.sum(m[:, None] * v ** 2)))
    Nf = 3.0 * (N - 1.0)
    kT = U.k_boltz * T
    q = np.ones(M) * tau_T ** 2 * kT
    q[0] *= Nf
    p_eta = rs.normal(0.0, np.sqrt(q * kT))
    L = nside * a * U.angstrom
    return dict(data=data, pos=pos * U.angstrom, vel=v, types=types, mass=mass_t, box=L,
                Nf=Nf, T=T, q=q, p_eta=p_eta, N=N, counts=counts)


def make_driver(w, precision="fp64", lanes_per_atom=0):
    from mdelast_amd import get_potential
    pots, eams = get_potential.tables_from_setfl(w["data"], "quintic")
    th = mdm.Thermostat(w["T"], w["q"], w["p_eta"], w["Nf"])
    ba = mdm.Barostat(0.0, 1.0, 0.0, mdm.Thermostat(w["T"], [1.0], [0.0], 1.0))
    box = mdm.Box(np.diag([w["box"]] * 3))
    return mdm.MDDriver(w["pos"], w["vel"], mdm.PotentialVector(pots), mdm.EAMVector(eams),
                        mdm.MDDriver.NVT, mdm.MDDriver.THERMOSTAT_KINETIC, th, ba, box, [1e-10],
                        w["Nf"], w["mass"], w["types"], 1e-15, True, 1.2, 1e3,
                        precision=precision, lanes_per_atom=lanes_per_atom), (pots, eams)


def pass_bytes(N, pairs_rc, mid_entries, list_entries, mid_frac):
    """Algorithmic HBM bytes of one launch of each pair pass in this layout: each
    atom's own data read once and its results written once; neighbour positions,
    their F' and the tables are gathers of data other atoms read too (re-use).
    Both mixed passes scan the mid list (pairs within r_c + 0.15 A); pairs_rc is
    reported beside them.

    force pass (k_force_mx): fixed-point position 16 B, F'/F'' 8 B, F 8 B, list
      length 4 B read, forces 3 x 8 B and energy 8 B written = 68 B per atom,
      plus 4 B per mid-list entry;
    density pass (k_density_mx): fixed-point position 16 B and list length 4 B
      read; F 8 B, rho 8 B, F'/F'' 16 + 8 B and the sampled g, |h|^2 64 B
      written = 124 B per atom, plus 4 B per mid-list entry scanned; on the
      fraction mid_frac of steps that refilter the mid list it reads the full
      list (4 B per entry), its length and the FP64 position (4 + 32 B) and
      writes the mid list (4 B per entry), its length and the reference
      position (4 + 32 B)."""
    force = 68.0 * N + 4.0 * mid_entries
    density = (124.0 * N + 4.0 * mid_entries
               + mid_frac * (4.0 * list_entries + 4.0 * mid_entries + 72.0 * N))
    return {"k_density_mx": density, "k_force_mx": force}


def step_bytes(N, list_entries, rebuild_every=20):
    """Algorithmic bytes of one whole step by the per-atom-step budget of the
    path's analysis (mixed precision): 238 B + 8 B per stored list entry (two
    passes) + (4 B per entry + 32 B) per rebuild, amortised."""
    K = list_entries / N
    return N * (238.0 + 8.0 * K + (4.0 * K + 32.0) / rebuild_every)


def read_pmc_profile():
    """Per-kernel counter figures from the committed rocprofv3 --pmc summary:
    {kernel: {hbm_bytes_per_dispatch, valu_busy, ta_busy, tcp_accesses_per_dispatch}}."""
    path = os.path.join(ROOT, "profiles", "pair_passes_nialco108k_pmc.txt")
    if not os.path.exists(path):
        return {}
    out = {}
    for line in open(path):
        line = line.split("#", 1)[0].strip()
        if ":" not in line:
            continue
        k, v = line.split(":", 1)
        if "." not in k:
            continue
        kern, field = k.strip().split(".", 1)
        try:
            out.setdefault(kern, {})[field] = float(v)
        except ValueError:
            pass
    return out


def cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cores_available": avail}


def bench_replicas(args):
    """Aggregate atom-steps/s of independent NVT replicas of uneven size,
    composition and temperature, advanced together (one host call per stretch)."""
    from mdelast_amd import get_potential, replicas as R
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = R.batched_replica_specs(args.replicas, edge=tuple(args.replica_edge))
    # one lane count for the whole batch (16 measured best for this table)
    lpa = args.lanes_per_atom or (16 if args.batched else 0)
    drivers = [R.make_replica_driver(s, pots, eams, masses, precision=args.precision,
                                     lanes_per_atom=lpa) for s in specs]
    graph_steps = args.graph_steps
    for d in drivers:
        d.graph_steps = graph_steps
    natoms = sum(s["N"] for s in specs)
    R.run_replicas(drivers, args.warmup, 1, 20, records=False, batched=args.batched)
    t0 = time.perf_counter()
    recs = R.run_replicas(drivers, args.steps, 1, 20, records=True, batched=args.batched)
    wall = time.perf_counter() - t0
    Tk = [float(r[-1][11]) for r in recs]
    out = {
        "metric": "atom-steps/s",
        "value": natoms * args.steps / wall,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{len(specs)} independent NVT replicas, ZnCu EAM, FCC edge "
                        f"{args.replica_edge[0]}-{args.replica_edge[1]} cells "
                        f"({4 * args.replica_edge[0] ** 3:,}-{4 * args.replica_edge[1] ** 3:,} "
                        "atoms), pure Cu or Cu-Zn up to 10 at.% Zn, T 100-900 K, "
                        "dt 1 fs, NHC M=10, Born every step, list r_skin 1.2 rebuilt every 20",
            "replicas": len(specs),
            "atoms_total": natoms,
            "atoms_min": min(s["N"] for s in specs),
            "atoms_max": max(s["N"] for s in specs),
            "precision": args.precision,
            "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default"),
            "graph_steps": graph_steps,
            "launches": "one per pass over all replicas" if args.batched
                        else "per replica, each on its own stream",
            "lanes_per_atom": lpa or "autotuned",
            "T_kin_final_minmax": [min(Tk), max(Tk)],
        },
    }
    print(json.dumps(out), flush=True)


def copy_calibration(mib):
    """hipMemcpy device-to-device of a buffer of mib MiB, three times: the HIP
    runtime's copy kernel reads and writes exactly that many bytes per launch,
    so a profiler's byte counters for it give their scale factor."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib * (1 << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipMemset(dst, 0, ctypes.c_size_t(nbytes))
    for _ in range(3):
        hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    return {"kernel": "HIP runtime device-to-device copy", "bytes_read": nbytes,
            "bytes_written": nbytes, "launches": 3}


def _pin_last_core():
    try:
        cpus = sorted(os.sched_getaffinity(0))
        os.sched_setaffinity(0, {cpus[-1]})
        return cpus
    except (AttributeError, OSError):
        return None


def cpu_calibration(nsteps=300, rebuild_every=20):
    """The oracle at the settings of the reference's published step time
    (README.md:277-283: 500 atoms of Cu with 30 Zn, quintic tables, list r_skin 1.2
    rebuilt every 20 steps, observables every step; 3.8 ms/step on one core of a
    Xeon Gold 6534, GCC 14.2 -O2): NPT at 300 K, dt 2 fs, chains of 10, from the
    FCC lattice a = 3.6345 A with Zn placed by RandomState(42)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    data = setfl.read_setfl("ZnCu.eam.alloy")
    a, side, T = 3.6345, 5, 300.0
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    idx = np.indices((side, side, side)).reshape(3, -1).T
    pos = ((idx[:, None, :] + 0.25 + basis[None, :, :]) * a).reshape(-1, 3) - side * a / 2.0
    N = len(pos)
    types = np.full(N, data.names.index("Cu"), dtype=np.uint32)
    types[np.random.RandomState(42).choice(N, 30, replace=False)] = data.names.index("Zn")
    mass_t = np.array([e.atomic_weight / U.Navogadro for e in data.elements])
    m = mass_t[types]
    rs = np.random.RandomState(0)
    sig = np.sqrt(U.k_boltz * T / m)
    v = np.stack([rs.normal(0.0, sig) for _ in range(3)], axis=1)
    v -= (v * m[:, None]).sum(0) / m.sum()
    v *= np.sqrt((3.0 * N - 3.0) * U.k_boltz * T / 2.0 / (0.5 * np.sum(m[:, None] * v ** 2)))
    Nf, kT = 3.0 * (N - 1.0), U.k_boltz * T
    q = np.ones(10) * (5e-14) ** 2 * kT
    q[0] *= Nf
    qb = np.ones(10) * (5e-13) ** 2 * kT
    W = Nf * (5e-13) ** 2 * kT
    o = O.MD(positions=pos * U.angstrom, velocities=v, types=types, mass=mass_t,
             box=[side * a * U.angstrom] * 3, data=data, ensemble=O.NPT,
             thermostat=(T, q, rs.normal(0.0, np.sqrt(q * kT)), Nf),
             barostat=(1e-10 * U.GPa, W, rs.normal(0.0, np.sqrt(W * kT))),
             baro_thermostat=(T, qb, rs.normal(0.0, np.sqrt(qb * kT)), 1.0), Nf=Nf, dt=2e-15,
             neighbors_list=True, r_skin=1.2)
    t0 = time.perf_counter()
    for k in range(nsteps):
        if k % rebuild_every == 0:
            o.calculate_neighbors()
        o.integrate()
    dt_step = (time.perf_counter() - t0) / nsteps
    return {"ms_per_step": 1e3 * dt_step, "published_ms_per_step": 3.8,
            "ratio_to_published": 1e3 * dt_step / 3.8,
            "sample": f"{nsteps} NPT steps of the 500-atom 6 at.% Zn cell, list 1.2 rebuilt "
                      f"every {rebuild_every}, observables every step"}


def cpu_baseline(w, nsteps=10, rebuild_every=20):
    """The oracle (single-threaded FP64 C restatement of the reference, gcc -O2) on one
    pinned core of this host: one construction of the same workload (its O(N^2) list
    build and forces), then nsteps NVT steps with the list rebuilt every
    rebuild_every steps (nsteps >= 100 gives >= 5 rebuilds). Per step = mean step
    time + construction time / rebuild_every when no rebuild fell in the window."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cpus = _pin_last_core()
    try:
        th = (w["T"], w["q"], w["p_eta"], w["Nf"])
        t0 = time.perf_counter()
        o = O.MD(positions=w["pos"], velocities=w["vel"], types=w["types"], mass=w["mass"],
                 box=[w["box"]] * 3, data=w["data"], ensemble=O.NVT, thermostat=th,
                 barostat=(0.0, 1.0, 0.0), baro_thermostat=(w["T"], [1.0], [0.0], 1.0),
                 Nf=w["Nf"], dt=1e-15, neighbors_list=True, r_skin=1.2)
        t1 = time.perf_counter()
        nreb = 0
        for k in range(nsteps):
            if k > 0 and k % rebuild_every == 0:
                o.calculate_neighbors()
                nreb += 1
            o.integrate()
        t2 = time.perf_counter()
        calib = cpu_calibration()
    finally:
        if cpus:
            os.sched_setaffinity(0, set(cpus))
    if nreb > 0:
        per_step = (t2 - t1) / nsteps
        how = f"mean over {nsteps} steps with {nreb} rebuilds in the window"
    else:
        t_step = (t2 - t1) / nsteps
        per_step = t_step + (t1 - t0) / rebuild_every
        how = (f"step mean + construction/{rebuild_every} (no rebuild in the "
               f"{nsteps}-step window; the construction stands in for one)")
    out = {"value": w["N"] / per_step, "unit": "atom-steps/s", "cores": 1, "kind": "port",
           "sample": (f"oracle (FP64 C restatement, gcc -O2), 1 pinned core, same "
                      f"{w['N']}-atom workload: construction with O(N^2) list build "
                      f"{t1 - t0:.1f} s, {nsteps} NVT steps {t2 - t1:.1f} s; per step = {how}; "
                      f"50,000 steps would take {50000 * per_step / 3600:.1f} h (extrapolated)"),
           "ms_per_step": 1e3 * per_step, "calibration": calib}
    out.update(cpu_info())
    return out


def learned_cv(n_eval):
    """The learned control-variate comparison in a child process (`--workload cv`):
    the control variate uses torch, which must bind the HIP runtime before the
    library does, while this process's plain MD path runs on the system runtime."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "cv",
           "--cv-eval-frames", str(n_eval)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    if res.returncode != 0 or not lines:
        return {"error": f"child exited {res.returncode}: {res.stderr.strip()[-400:]}"}
    return json.loads(lines[-1]).get("learned_cv")


def _learned_cv_here(n_eval):
    """GPU-seconds to C11, C12 and C44 at 2 SE <= 1 GPa, plain and with the
    learned Stein control variate, as a production run would use it: 4,000-atom
    alpha-brass (6 at.% Zn random fcc, 10^3 cells, a = 3.667 A), NVT 300 K, dt
    1 fs, frames every 20 steps; three fresh alloy samples (placement seeds)
    x three fresh trajectories each; in every run the plain estimate and the
    cross-fitted controlled estimate of the same frames (no training run: each
    half of a run is controlled by coefficients fitted on the other half, and
    the fitting is timed into the controlled route's cost)."""
    from mdelast_amd import cv
    try:
        r = cv.brass_crossfit_study(
            n_frames=n_eval, log=lambda m: print(m, file=sys.stderr, flush=True))
    except Exception as e:        # reported, never hides the throughput line
        return {"error": f"{type(e).__name__}: {e}"}

    def f(x):
        return [round(float(v), 6) for v in np.ravel(x)]

    runs = r["runs"]
    return {
        "workload": "CuZn 6 at.% Zn random fcc 10^3 cells, 4000 atoms (240 Zn), a=3.667 A, "
                    "NVT 300 K, dt 1 fs, NHC M=10, frames every 20 steps; "
                    f"{len(r['state_seeds'])} fresh alloy samples x {r['runs_per_state']} "
                    f"fresh trajectories of {r['n_frames']} frames",
        "protocol": "cross-fitted (two halves, each controlled by a fit on the other), "
                    "no training run; fits and captures timed into the controlled cost",
        "heads": ["C11", "C12", "C44"],
        "gpu_s_plain": f(r["gpu_s_plain"]),
        "gpu_s_cv": f(r["gpu_s_cv"]),
        "gpu_s_plain_worst": round(r["gpu_s_plain_worst"], 6),
        "gpu_s_cv_worst": round(r["gpu_s_cv_worst"], 6),
        "speedup": round(r["speedup"], 4),
        "training_gpu_s": 0.0,
        "cost_plain_us_per_frame": round(1e6 * r["cost_plain_frame_s"], 2),
        "cost_cv_us_per_frame": round(1e6 * r["cost_cv_frame_s"], 2),
        "cost_cv_capture_us_per_frame": round(
            1e6 * float(np.mean([x["t_capture_frame"] for x in runs])), 2),
        "cost_cv_fits_us_per_frame": round(
            1e6 * float(np.mean([x["t_fit_frame"] for x in runs])), 2),
        "C_plain_GPa": [f(x["C_plain"]) for x in runs],
        "C_cv_GPa": [f(x["C_cv"]) for x in runs],
        "se_plain_GPa": f(r["se_plain"]),
        "se_cv_GPa": f(r["se_cv"]),
        "mean_control_over_se": [f(x["mean_s"] / x["se_mean_s"]) for x in runs],
        "control_z_rms": f(r["control_z_rms"]),
        "eval_var_ratio": f(np.mean([x["var_ratio"] for x in runs], 0)),
        "eval_se2_ratio": f((np.asarray(r["se_cv"]) / np.asarray(r["se_plain"])) ** 2),
        "enabled_per_fold": [x["enabled"] for x in runs],
        "fields": "displacement, energy- and pressure-proxy-weighted displacement, and the "
                  "zero-variance fields of a fitted stress proxy (grad h_a, (h_a - h0_a) grad h_b)",
        "uses_stress_proxy_fields": True,
        "uses_network_fields": [False, False, False],
        "eval_runs": len(runs),
        "nblocks": r["nblocks"],
        "t_step_us": round(r["t_step_s"] * 1e6, 3),
        "frames_eval": r["n_frames"] * len(runs),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--nside", type=int, default=30)
    ap.add_argument("--lanes-per-atom", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-steps", type=int, default=10,
                    help="NVT steps of the oracle after its construction (>= 100 for a "
                         "window with >= 5 list rebuilds; about 0.5 s each on one core)")
    ap.add_argument("--workload", default="nialco108k", choices=["nialco108k", "replicas", "cv"],
                    help="nialco108k: the single 108,000-atom instance (default); replicas: "
                         "independent NVT Cu / Cu-Zn replicas of 500-32,000 atoms advanced side "
                         "by side; cv: only the learned control-variate comparison")
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--replica-edge", type=int, nargs=2, default=[5, 20], metavar=("LO", "HI"),
                    help="replicas: FCC edges (cells) drawn from LO..HI (4 edge^3 atoms)")
    ap.add_argument("--graph-steps", type=int, default=20,
                    help="replay aligned chunks of this many steps as one captured HIP graph "
                         "(the single system; per replica, or one graph of batched steps with "
                         "--batched; 0: launch step by step)")
    ap.add_argument("--batched", action="store_true",
                    help="replicas: each pass of a step as one launch over every replica "
                         "(mdl_replicas_run) instead of per-replica launches")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="hardware queues of the process (GPU_MAX_HW_QUEUES, at most 32)")
    ap.add_argument("--no-learned-cv", action="store_true",
                    help="skip the learned control-variate comparison (4,000-atom alpha-brass)")
    ap.add_argument("--cv-eval-frames", type=int, default=3000,
                    help="frames per run of the control-variate comparison (9 runs)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB three times "
                         "(a kernel of known HBM traffic, to scale rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; the workload is a single "
              "instance on one MI355X (multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)
    mdm.VERBOSE = False
    if args.hw_queues:
        # read by the HIP runtime when it initialises (before the library loads)
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(1, args.hw_queues)))
    if args.workload == "replicas":
        return bench_replicas(args)
    if args.workload == "cv":
        r = _learned_cv_here(args.cv_eval_frames)
        print(json.dumps({"metric": "GPU-s to C11/C12/C44 at +/-1 GPa (worst head)",
                          "value": r.get("gpu_s_cv_worst"), "unit": "s",
                          "higher_is_better": False, "n_gpus": 1, "learned_cv": r}), flush=True)
        return
    w = build_workload(nside=args.nside)
    d, keep = make_driver(w, args.precision, args.lanes_per_atom)
    if args.graph_steps > 0:
        d.graph_steps = args.graph_steps
    d.run(args.warmup, 1, 20, records=False)
    mid0 = d.mid_list_builds
    t0 = time.perf_counter()
    rec = d.run(args.steps, 1, 20, records=True)   # synchronises before returning
    t1 = time.perf_counter()
    wall = t1 - t0
    kt = d.kernel_times()
    N = w["N"]
    value = N * args.steps / wall
    entries = d.neighbor_list_entries
    pairs_rc = float(d.pairs_within_rc().sum())
    mid_entries = float(d.mid_list_lengths().sum())
    mid_frac = (d.mid_list_builds - mid0) / args.steps
    alg = pass_bytes(N, pairs_rc, mid_entries, entries, mid_frac)
    times = {"k_density_mx": kt["density_ms"], "k_force_mx": kt["force_ms"]}
    # counter figures are committed for the mixed-precision kernels only
    pmc = read_pmc_profile() if args.precision == "mixed" else {}

    def kernel_roof(k):
        t = times[k] * 1e-3
        r = {"algorithmic_bytes_per_launch": alg[k], "ms_per_launch": times[k],
             "achieved": alg[k] / t / 1e9}
        r["frac"] = r["achieved"] / HBM_PEAK_GBS
        c = pmc.get(k, {})
        traffic = c.get("hbm_bytes_per_dispatch")
        r["traffic"] = traffic
        r["achieved_counter"] = traffic / t / 1e9 if traffic else None
        r["frac_counter"] = r["achieved_counter"] / HBM_PEAK_GBS if traffic else None
        for f in ("valu_busy", "ta_busy", "tcp_accesses_per_pair"):
            r[f] = c.get(f)
        return r

    dom = max(times, key=times.get)
    other = "k_force_mx" if dom == "k_density_mx" else "k_density_mx"
    rd, ro = kernel_roof(dom), kernel_roof(other)
    sb = step_bytes(N, entries)
    out = {
        "metric": "atom-steps/s",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"NiAlCo fcc {args.nside}^3 cells, {N} atoms "
                        f"({w['counts'][0]} Ni/{w['counts'][1]} Al/{w['counts'][2]} Co), "
                        "a=3.56 A, NVT 600 K, dt 1 fs, NHC M=10, Born every step, "
                        "list r_skin 1.2 rebuilt every 20 steps",
            "atoms": N,
            "ns_per_day": 1e-6 * (args.steps / wall) * 86400.0,
            "precision": args.precision,
            "neighbor_entries_per_atom": entries / N,
            "pairs_within_rc_per_atom": pairs_rc / N,
            "mid_list_entries_per_atom": mid_entries / N,
            "mid_list_filterings_per_step": mid_frac,
            "list_builds": d.neighbor_list_builds,
            "lanes_per_atom": d.lanes_per_atom_used,
            "graph_steps": args.graph_steps,
            "nhc_general_path": d.nhc_general_path_count,
            "kernel_ms": {k: round(v, 5) for k, v in kt.items()},
            "final_T_kin": float(rec[-1][11]),
            "final_etot_cons_err": float(rec[-1][7]),
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": rd["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": rd["frac"],
            "traffic": rd["traffic"],
            "algorithmic_bytes_per_launch": rd["algorithmic_bytes_per_launch"],
            "achieved_counter": rd["achieved_counter"],
            "frac_counter": rd["frac_counter"],
            "valu_busy": rd["valu_busy"],
            "ta_busy": rd["ta_busy"],
            "tcp_accesses_per_pair": rd["tcp_accesses_per_pair"],
            "ms_per_launch": rd["ms_per_launch"],
            other: ro,
            "step": {"bytes_per_atom_step": sb / N, "achieved": sb * args.steps / wall / 1e9,
                     "frac": sb * args.steps / wall / 1e9 / HBM_PEAK_GBS},
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline and int(os.environ.get("RANK", "0")) == 0:
        out["cpu_baseline"] = cpu_baseline(w, nsteps=args.cpu_baseline_steps)
    if args.copy_calibration_mib > 0:
        out["copy_calibration"] = copy_calibration(args.copy_calibration_mib)
    if not args.no_learned_cv:
        del d, keep
        out["learned_cv"] = learned_cv(args.cv_eval_frames)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
