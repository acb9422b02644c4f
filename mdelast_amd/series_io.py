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

``summarize_run`` reduces a series file to the benchmark quantities through a
table (quantity, unit, reducer), and ``check_value`` / ``verify_md`` hold them
to the entries of the reference's expected-value files
(``benchmarks/expected/*.json``: band, bounds or report only), so those files
apply to GPU-produced output unchanged.
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


# ---------------------------------------------------------------------------
# The benchmark quantities of one run, as a table: each quantity names its
# unit and how it reduces the series file. The quantities, their definitions
# and the acceptance rules are those of the reference's expected-value files
# (benchmarks/expected/*.json, documented with benchmarks/bench_lib.py:47-157
# and benchmarks/verify.py:80-110); the code is this package's own.
# ---------------------------------------------------------------------------
class RunView:
    """A series file seen as one run: production starts at the stored sample
    index ``n0`` (series/temperature_kin/cycles_thermalization); one sample is
    stored every ``stride`` steps (thermal_results/sample_stride, else 1)."""

    def __init__(self, f):
        self.f = f
        self.n0 = int(f["series/temperature_kin/cycles_thermalization"])
        self.stride = int(f["thermal_results/sample_stride"]) \
            if "thermal_results/sample_stride" in f else 1
        self.samples = int(len(f["series/time"]))

    def result(self, key, default=None):
        k = f"thermal_results/{key}"
        return np.asarray(self.f[k]) if k in self.f else default

    def mean(self, name):
        """The run's stored production mean of a scalar series."""
        return float(self.f[f"series/{name}/mean"])

    def samples_of(self, name):
        return np.asarray(self.f[f"series/{name}/data"], dtype=float)

    def voigt_mean(self, prefix, group):
        """Mean over the Voigt entries of a cubic group (11: the three normal
        diagonals, 12: the three normal off-diagonals, 44: the three shears) of a
        matrix series' production means."""
        return float(np.mean([self.mean(f"{prefix}_{a}_{b}") for a, b in CUBIC[group]]))

    def econs_relative(self):
        e = self.samples_of("etot_cons")
        return e / e[0] - 1.0


def _econs_drift_per_ns(v):
    """Least-squares slope of E_cons / E_cons(0) - 1 against time over production."""
    t = np.asarray(v.f["series/time"], dtype=float)[v.n0:] * 1e9
    return float(np.polyfit(t, v.econs_relative()[v.n0:], 1)[0])


def _quantity_table():
    """[(name, unit, reducer(RunView) -> float)]"""
    table = [(f"C{g}", "GPa", lambda v, g=g: float(v.result(f"elasticity_tensor_{g}")) / GPA)
             for g in CUBIC]
    table.append(("cubic_spread_C11", "GPa",
                  lambda v: float(np.ptp(np.diag(v.result("elasticity_tensor"))[:3])) / GPA))
    table += [(f"CB{g}", "GPa", lambda v, g=g: v.mean(f"born_elasticity_tensor_{g}") / GPA)
              for g in CUBIC]
    table += [(f"fluc{g}", "GPa", lambda v, g=g: v.voigt_mean("stress_fluc", g) / GPA)
              for g in CUBIC]
    table += [(f"kin{g}", "GPa", lambda v, g=g: v.voigt_mean("elasticity_ideal_gas", g) / GPA)
              for g in ("11", "44")]
    table += [
        ("K_npt", "GPa", lambda v: float(v.result("bulk_modulus", 0.0)) / GPA),
        ("T_kin", "K", lambda v: v.mean("temperature_kin")),
        # 1 / <1/T_conf> (the estimator's form) and <T_conf> per sample
        ("T_conf_recip", "K", lambda v: 1.0 / v.mean("temperature_conf_inverse")),
        ("T_conf_sample", "K",
         lambda v: float(np.mean(1.0 / v.samples_of("temperature_conf_inverse")[v.n0:]))),
        ("pressure", "GPa", lambda v: v.mean("pressure") / GPA),
        ("L_L0", "", lambda v: v.mean("L_L0")),
        ("V_V0", "", lambda v: v.mean("V_V0")),
        # over the whole run, thermalisation included
        ("econs_err_max", "", lambda v: float(np.max(np.abs(v.econs_relative())))),
        ("econs_drift_per_ns", "1/ns", _econs_drift_per_ns),
    ]
    return table


QUANTITIES = _quantity_table()


def summarize_run(path) -> dict:
    """Every quantity of ``QUANTITIES`` for the run in series file ``path`` (GPa,
    K, 1/ns, ratios), plus its length: n_steps and n_thermalization in MD steps
    and the sample stride."""
    f = SeriesFile(path)
    try:
        v = RunView(f)
        out = {"n_steps": v.samples * v.stride, "n_thermalization": v.n0 * v.stride,
               "sample_stride": v.stride}
        for name, _unit, reduce in QUANTITIES:
            out[name] = reduce(v)
    finally:
        f.close()
    return out


# ---------------------------------------------------------------------------
# Acceptance rules of an expected-value entry: a band (expected +/- tol),
# bounds (min and/or max), or a value only reported (info).
# ---------------------------------------------------------------------------
def _rule(spec):
    if spec.get("info"):
        return "info", None
    if "expected" in spec:
        return "band", (float(spec["expected"]), float(spec["tol"]))
    return "bounds", (float(spec.get("min", -math.inf)), float(spec.get("max", math.inf)))


def check_value(name, actual, spec):
    """(status, message) of one value under its expected-value entry: PASS or FAIL
    for a band or bounds, INFO for a report-only entry; a missing or non-finite
    value fails."""
    unit = spec.get("unit", "")
    if actual is None or not math.isfinite(float(actual)):
        return "FAIL", f"{name}: no finite value"
    kind, arg = _rule(spec)
    shown = f"{name} = {float(actual):.6g}{(' ' + unit) if unit else ''}"
    if kind == "info":
        return "INFO", f"{shown}  [reported only: {spec['info']}]"
    if kind == "band":
        centre, tol = arg
        off = float(actual) - centre
        ok = abs(off) <= tol
        return ("PASS" if ok else "FAIL"), f"{shown}  [{centre:.6g} +/- {tol:.3g}, off by {off:+.3g}]"
    lo, hi = arg
    ok = lo <= float(actual) <= hi
    return ("PASS" if ok else "FAIL"), f"{shown}  [within {lo:.3g} .. {hi:.3g}]"


def verify_md(expected: dict, stage_paths: dict, require_full_length: bool = True):
    """Check the MD stages of an expected-value file against series files
    (``stage_paths`` = {"npt": path, "nvt": path}). Returns (rows, summaries):
    one (status, message) row per checked entry, in the file's order, and each
    stage's summarize_run. A stage without a file fails; so does a run whose
    length differs from the reference's when ``require_full_length`` (the
    tolerances are those of full-length runs). The E_cons drift is checked by
    magnitude (econs_drift_abs_per_ns). With a zero_kelvin_reference block, the
    NVT constants are also held to its relative deviation."""
    rows, summaries = [], {}
    for stage, st in expected["stages"].items():
        path = stage_paths.get(stage)
        if path is None:
            rows.append(("FAIL", f"{stage}: no series file"))
            continue
        got = summarize_run(path)
        summaries[stage] = got
        want = (st["n_steps"], st["n_thermalization"])
        have = (got["n_steps"], got["n_thermalization"])
        if require_full_length and have != want:
            rows.append(("FAIL", f"{stage}: run length {have[1]} + {have[0] - have[1]} steps "
                                 f"(thermalisation + production), the reference's is "
                                 f"{want[1]} + {want[0] - want[1]}"))
        values = dict(got, econs_drift_abs_per_ns=abs(got["econs_drift_per_ns"]))
        rows += [check_value(f"{stage}.{q}", values.get(q), spec)
                 for q, spec in st["quantities"].items()]
    zk = expected.get("zero_kelvin_reference")
    if zk and "nvt" in summaries:
        dev = zk["max_rel_deviation"]
        for q, ref in zk["values_GPa"].items():
            rel = (summaries["nvt"][q] - ref) / ref
            rows.append(check_value(f"nvt.{q} relative to the 0 K value {ref} GPa", rel,
                                    {"min": -dev, "max": dev}))
    return rows, summaries
