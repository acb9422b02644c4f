"""Series output with the reference's ``series.h5`` schema, and the benchmark
summary and checks that read it.

The reference writes ``series.h5`` after every thermal average
(``MolecularDynamics.save_h5``, molecular_dynamics.py:1182-1195; layout from
``TimeSeries.write_h5`` and friends, series.py): ``thermal_results/<name>``,
``series/time``, and per series ``series/<name>/{data, mean, std, std_mean,
mean_cum, std_cum, cycles_thermalization}``, vectors as ``<name>_<i>`` and
matrices as ``<name>_<i>_<j>`` (1-based). h5py is not part of this stack, so the
same keys go into one ``series.npz`` (``numpy.savez``; ``numpy.load`` reads them
back with the same paths). The Gaussian plotting aids ``data_gauss`` and
``gauss`` are not written.

``summarize_run`` and ``check_value`` follow ``benchmarks/bench_lib.py:47-151``,
``verify_md`` the MD part of ``benchmarks/verify.py:80-110``, so the reference's
expected-value files (``benchmarks/expected/*.json``) apply to GPU-produced
output unchanged.
"""

from __future__ import annotations

import math

import numpy as np

GPA = 1e10  # dyn/cm^2 per GPa

STATS = ("data", "mean", "std", "std_mean", "mean_cum", "std_cum", "cycles_thermalization")

# Voigt index groups of the cubic averages (bench_lib.py:24-28)
CUBIC = {"11": ((1, 1), (2, 2), (3, 3)), "12": ((1, 2), (1, 3), (2, 3)),
         "44": ((4, 4), (5, 5), (6, 6))}


def _series_items(series):
    """(key, TimeSeries) for every scalar series, vectors and matrices expanded."""
    from .molecular_dynamics import TimeSeries, TimeSeriesMatrix, TimeSeriesVector
    for name, value in vars(series).items():
        if isinstance(value, TimeSeries):
            yield name, value
        elif isinstance(value, TimeSeriesVector):
            for i, ts in enumerate(value.data):
                yield f"{name}_{i + 1}", ts
        elif isinstance(value, TimeSeriesMatrix):
            for i, row in enumerate(value.data):
                for j, ts in enumerate(row):
                    yield f"{name}_{i + 1}_{j + 1}", ts


def save_series(series, path) -> str:
    """Write ``series`` (a ``molecular_dynamics.Series``) to ``path`` (.npz)."""
    out = {}
    for key, value in vars(series.thermal_results).items():
        a = np.asarray(value) if value is not None else None
        if a is None or a.dtype == object:
            continue        # only numeric results have a place in the file
        out[f"thermal_results/{key}"] = a
    out["series/time"] = np.asarray(series.time, dtype=float)
    for name, ts in _series_items(series):
        for stat in STATS:
            v = getattr(ts, stat, None)
            if v is not None:
                out[f"series/{name}/{stat}"] = np.asarray(v)
    path = str(path)
    np.savez(path, **out)
    return path if path.endswith(".npz") else path + ".npz"


class SeriesFile:
    """Read access by the series.h5 paths: ``f["series/pressure/mean"]``."""

    def __init__(self, path):
        self._z = np.load(str(path), allow_pickle=False)

    def __getitem__(self, key):
        return self._z[key]

    def __contains__(self, key):
        return key in self._z.files

    def close(self):
        self._z.close()


def _mean(f, name):
    return float(f[f"series/{name}/mean"])


def _data(f, name):
    return np.asarray(f[f"series/{name}/data"], dtype=float)


def _cubic_mean(f, prefix):
    return {key: float(np.mean([_mean(f, f"{prefix}_{i}_{j}") for i, j in comps]))
            for key, comps in CUBIC.items()}


def summarize_run(path) -> dict:
    """The benchmark quantities of one run (bench_lib.py:47-107), in GPa, K, 1/ns
    and dimensionless ratios: C11, C12, C44, CB*, fluc*, kin11, kin44, K_npt, T_kin,
    T_conf_recip, T_conf_sample, pressure, L_L0, V_V0, econs_err_max,
    econs_drift_per_ns, cubic_spread_C11, and the run length."""
    f = SeriesFile(path)
    try:
        n0 = int(f["series/temperature_kin/cycles_thermalization"])
        stride = int(f["thermal_results/sample_stride"]) \
            if "thermal_results/sample_stride" in f else 1
        out = {"n_steps": int(len(f["series/time"])) * stride,
               "n_thermalization": n0 * stride, "sample_stride": stride,
               "C11": float(f["thermal_results/elasticity_tensor_11"]) / GPA,
               "C12": float(f["thermal_results/elasticity_tensor_12"]) / GPA,
               "C44": float(f["thermal_results/elasticity_tensor_44"]) / GPA}
        full = np.asarray(f["thermal_results/elasticity_tensor"]) / GPA
        out["cubic_spread_C11"] = float(np.ptp(np.diag(full)[:3]))
        for key in CUBIC:
            out[f"CB{key}"] = _mean(f, f"born_elasticity_tensor_{key}") / GPA
        fluc = _cubic_mean(f, "stress_fluc")
        kin = _cubic_mean(f, "elasticity_ideal_gas")
        for key in CUBIC:
            out[f"fluc{key}"] = fluc[key] / GPA
        out["kin11"] = kin["11"] / GPA
        out["kin44"] = kin["44"] / GPA
        out["K_npt"] = float(f["thermal_results/bulk_modulus"]) / GPA \
            if "thermal_results/bulk_modulus" in f else 0.0
        out["T_kin"] = _mean(f, "temperature_kin")
        inv = _data(f, "temperature_conf_inverse")
        out["T_conf_recip"] = 1.0 / _mean(f, "temperature_conf_inverse")
        out["T_conf_sample"] = float(np.mean(1.0 / inv[n0:]))
        out["pressure"] = _mean(f, "pressure") / GPA
        out["L_L0"] = _mean(f, "L_L0")
        out["V_V0"] = _mean(f, "V_V0")
        econs = _data(f, "etot_cons")
        rel = econs / econs[0] - 1.0
        out["econs_err_max"] = float(np.max(np.abs(rel)))
        t_ns = np.asarray(f["series/time"], dtype=float)[n0:] * 1e9
        out["econs_drift_per_ns"] = float(np.polyfit(t_ns, rel[n0:], 1)[0])
    finally:
        f.close()
    return out


def check_value(name, actual, spec):
    """(status, message) of one value against its spec (bench_lib.py:117-151):
    ``expected`` and ``tol``, or ``min``/``max`` bounds, or ``info`` (report only)."""
    if actual is None or (isinstance(actual, float) and not math.isfinite(actual)):
        return "FAIL", f"{name}: missing or non-finite value"
    unit = spec.get("unit", "")
    if spec.get("info"):
        return "INFO", f"{name} = {actual:.6g} {unit} (not checked: {spec['info']})"
    if "expected" in spec:
        exp, tol = float(spec["expected"]), float(spec["tol"])
        d = actual - exp
        return ("PASS" if abs(d) <= tol else "FAIL"), \
            f"{name} = {actual:.6g} {unit}, expected {exp:.6g} +/- {tol:.3g} (diff {d:+.3g})"
    lo, hi = spec.get("min", -math.inf), spec.get("max", math.inf)
    return ("PASS" if lo <= actual <= hi else "FAIL"), \
        f"{name} = {actual:.6g} {unit}, required [{lo:.3g}, {hi:.3g}]"


def verify_md(expected: dict, stage_paths: dict, require_full_length: bool = True):
    """Rows (status, message) of an expected-value file's MD stages
    (benchmarks/verify.py:80-110) against series files, ``stage_paths`` =
    {"npt": path, "nvt": path}. A run length different from the reference's is a
    FAIL when ``require_full_length`` (the tolerances hold for full-length runs)."""
    rows, summaries = [], {}
    for stage, st in expected["stages"].items():
        if stage not in stage_paths:
            rows.append(("FAIL", f"{stage}: no series file"))
            continue
        s = summarize_run(stage_paths[stage])
        summaries[stage] = s
        if require_full_length and \
                (s["n_steps"], s["n_thermalization"]) != (st["n_steps"], st["n_thermalization"]):
            rows.append(("FAIL", f"{stage}: run length {s['n_thermalization']}+"
                                 f"{s['n_steps'] - s['n_thermalization']} differs from the "
                                 f"reference's {st['n_thermalization']}+"
                                 f"{st['n_steps'] - st['n_thermalization']}"))
        s = dict(s, econs_drift_abs_per_ns=abs(s["econs_drift_per_ns"]))
        for q, spec in st["quantities"].items():
            rows.append(check_value(f"{stage}.{q}", s.get(q), spec))
    zk = expected.get("zero_kelvin_reference")
    if zk and "nvt" in summaries:
        for q, ref in zk["values_GPa"].items():
            dev = (summaries["nvt"][q] - ref) / ref
            rows.append(check_value(f"nvt.{q} vs 0 K reference {ref} GPa", dev,
                                    {"min": -zk["max_rel_deviation"],
                                     "max": zk["max_rel_deviation"]}))
    return rows, summaries
