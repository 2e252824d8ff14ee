"""Autotuner of the GPU launch shapes (no reference counterpart).

The report kernel's waves per CU at the 64- and 128-diagonal windows, the
pebble pool of the 64-diagonal window and the batched aligner's waves per CU
at its 64-diagonal and its wider windows (MiTuning, include/marvl_align.h)
change how fast a pass runs, never what it writes.  The tuner times short calibration runs of the caller's own
workload with each candidate, keeps the fastest, and stores the choice per
device (architecture and compute units) in a JSON cache, so later processes
on the same kind of GPU load it instead of calibrating again.

    python -m marvel_amd.tune --show        # the cached choice for this GPU
"""

from __future__ import annotations

import ctypes as C
import json
import os
import time

from .lib import MiTuning, lib

FIELDS = ("report_waves_per_cu", "report_pool_log2", "report_l1_waves_per_cu", "batch_waves_per_cu",
          "batch_wide_waves_per_cu")
DEFAULT = {"report_waves_per_cu": 32, "report_pool_log2": 15, "report_l1_waves_per_cu": 16, "batch_waves_per_cu": 32,
           "batch_wide_waves_per_cu": 16}
REPORT_CANDIDATES = [(w, p) for w in (16, 24, 32) for p in (14, 15, 16)]
L1_CANDIDATES = (8, 16, 24, 32)
# (64-diagonal window, wider windows): the levels are timed apart, so each
# run scores one candidate of each and the two are chosen independently
BATCH_CANDIDATES = ((16, 8), (24, 12), (32, 16), (32, 20))


def cache_path() -> str:
    base = os.environ.get("MI_TUNE_CACHE") or os.path.join(
        os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache"), "marvel_amd")
    return os.path.join(base, "tuning.json")


def device_key() -> str:
    name = C.create_string_buffer(64)
    cus = C.c_int(0)
    if lib.mi_device_info(name, 64, C.byref(cus)) != 0:
        raise RuntimeError(lib.last_error())
    return f"{name.value.decode()}:{cus.value}cu"


def apply(choice: dict):
    t = MiTuning(*[int(choice.get(f, 0)) for f in FIELDS])
    if lib.mi_set_tuning(C.byref(t)) != 0:
        raise RuntimeError(lib.last_error())


def current() -> dict:
    t = MiTuning()
    lib.mi_get_tuning(C.byref(t))
    return {f: getattr(t, f) or DEFAULT[f] for f in FIELDS}


def load_cached() -> dict | None:
    try:
        with open(cache_path()) as f:
            return json.load(f).get(device_key())
    except (OSError, ValueError):
        return None


def save(choice: dict):
    p = cache_path()
    os.makedirs(os.path.dirname(p), exist_ok=True)
    try:
        with open(p) as f:
            allc = json.load(f)
    except (OSError, ValueError):
        allc = {}
    allc[device_key()] = choice
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(allc, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def _time_report(step, reps=1) -> float:
    """Report-kernel seconds (HIP events, both strands) of `step()`, the best of reps."""
    best = float("inf")
    for _ in range(reps):
        s1, s2 = step()
        best = min(best, s1["t_report_kernel"] + s2["t_report_kernel"])
    return best


class _FirstLevel:
    """Pins the report kernel's first window (MI_REPORT_FIRST_LEVEL) for a
    calibration loop, so the library's own choice of it does not move
    between candidates."""

    def __init__(self, level: int):
        self.level = str(level)

    def __enter__(self):
        self.prev = os.environ.get("MI_REPORT_FIRST_LEVEL")
        os.environ["MI_REPORT_FIRST_LEVEL"] = self.level

    def __exit__(self, *exc):
        if self.prev is None:
            del os.environ["MI_REPORT_FIRST_LEVEL"]
        else:
            os.environ["MI_REPORT_FIRST_LEVEL"] = self.prev
        return False


def tune_report(step) -> dict:
    """Picks the report kernel's shapes with `step()` (one block-pair
    comparison returning both passes' stats): waves per CU and pool at the
    64-diagonal window, then waves per CU at the 128-diagonal window.  The
    latter are timed with every pair started at the 128-diagonal window
    (MI_REPORT_FIRST_LEVEL=1): in an ordinary step that window re-runs only
    the few pairs that outgrow 64 diagonals, which no candidate changes
    measurably, while on repeat-rich blocks it carries as much work as the
    64-diagonal window and the shape matters (8 waves per CU ran 19 % slower
    than 16 to 32 there)."""
    base = current()
    timings = {}
    with _FirstLevel(0):
        for w, p in REPORT_CANDIDATES:
            apply({**base, "report_waves_per_cu": w, "report_pool_log2": p})
            timings[f"w{w}_p{p}"] = _time_report(step)
    w, p = min(REPORT_CANDIDATES, key=lambda c: timings[f"w{c[0]}_p{c[1]}"])
    best = {**base, "report_waves_per_cu": w, "report_pool_log2": p}
    with _FirstLevel(1):
        for l1 in L1_CANDIDATES:
            apply({**best, "report_l1_waves_per_cu": l1})
            timings[f"l1_{l1}"] = _time_report(step)
    best["report_l1_waves_per_cu"] = min(L1_CANDIDATES, key=lambda c: timings[f"l1_{c}"])
    apply(best)
    return {"choice": best, "timings_s": {k: round(v, 5) for k, v in timings.items()}}


def tune_batch(run) -> dict:
    """Picks the batched aligner's waves per CU with `run()` (one batch
    returning its work list: kernel microseconds per window level at [2:6]).
    The 64-diagonal level's time depends only on its own waves and the wider
    levels' only on theirs (the tasks each level re-runs do not depend on the
    launch shape), so one run per candidate pair scores both."""
    base = current()
    narrow, wide = {}, {}
    for w0, w1 in BATCH_CANDIDATES:
        apply({**base, "batch_waves_per_cu": w0, "batch_wide_waves_per_cu": w1})
        work = run()
        narrow[w0] = min(narrow.get(w0, float("inf")), work[2] / 1e6)
        wide[w1] = sum(work[3:6]) / 1e6
    best = {**base, "batch_waves_per_cu": min(narrow, key=narrow.get), "batch_wide_waves_per_cu": min(wide, key=wide.get)}
    apply(best)
    timings = {**{f"b{k}": v for k, v in narrow.items()}, **{f"wide{k}": v for k, v in wide.items()}}
    return {"choice": best, "timings_s": {k: round(v, 5) for k, v in timings.items()}}


def ensure(report_step=None, batch_run=None) -> dict:
    """The cached choice for this device when there is one; otherwise
    calibrates with the callables given, stores and applies the result."""
    cached = load_cached()
    if cached is not None:
        apply(cached)
        return {"source": "cache", "device": device_key(), "choice": cached}
    t0 = time.time()
    out = {"source": "calibrated", "device": device_key()}
    if report_step is not None:
        out["report"] = tune_report(report_step)
    if batch_run is not None:
        out["batch"] = tune_batch(batch_run)
    out["choice"] = current()
    out["calibration_s"] = round(time.time() - t0, 2)
    save(out["choice"])
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(prog="python -m marvel_amd.tune")
    ap.add_argument("--show", action="store_true")
    a = ap.parse_args()
    print(json.dumps({"device": device_key(), "cached": load_cached(), "cache": cache_path()}))
