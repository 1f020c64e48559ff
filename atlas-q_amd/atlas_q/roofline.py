"""Live per-kernel timing with HIP events for roofline accounting.

When enabled, instrumented launches record a start/end event pair on the
stream they are launched on (torch's current stream, which is the stream the
HIP entry points receive) together with the launch's algorithmic flops and
bytes.  ``summary()`` reports, per kernel, the mean launch duration and the
achieved rate against the MI355X peak of its bound.
"""
from __future__ import annotations

import contextlib
from collections import defaultdict

import torch

PEAK = {"fp64_mfma": 78.6, "fp64_valu": 78.6, "fp32_mfma": 157.3, "bf16_mfma": 2500.0,
        "hbm": 8000.0}
UNIT = {"fp64_mfma": "TFLOP/s", "fp64_valu": "TFLOP/s", "fp32_mfma": "TFLOP/s",
        "bf16_mfma": "TFLOP/s", "hbm": "GB/s"}

_enabled = False
_records = []


def enable(on: bool) -> None:
    global _enabled
    _enabled = on


def reset() -> None:
    _records.clear()


@contextlib.contextmanager
def timed(name: str, work: float, bound: str, nbytes: float = 0.0):
    """Time one launch.  work: algorithmic flops (MFMA / VALU bounds) or bytes
    (hbm); nbytes: the algorithmic bytes of a flop-counted launch, which make
    it HBM-bound when moving them at 8 TB/s takes longer than the flops at
    their peak."""
    if not _enabled:
        yield
        return
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    yield
    e.record()
    _records.append((name, s, e, float(work), bound, float(nbytes)))


class Span:
    """One logical call made of parts on several streams (for instance the
    same step of several batches that overlap): ``part()`` times each part on
    the current stream, and the record's duration runs from the earliest part
    start to the latest part end."""

    def __init__(self, name: str, bound: str):
        self.name, self.bound = name, bound
        self.starts, self.ends = [], []
        self.work = self.nbytes = 0.0

    @contextlib.contextmanager
    def part(self, work: float, nbytes: float = 0.0):
        if not _enabled:
            yield
            return
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        yield
        e.record()
        self.starts.append(s)
        self.ends.append(e)
        self.work += float(work)
        self.nbytes += float(nbytes)

    def close(self) -> None:
        if _enabled and self.starts:
            _records.append((self.name, self.starts, self.ends, self.work, self.bound,
                             self.nbytes))


def _elapsed_ms(s, e) -> float:
    """Event-pair duration; for a span, earliest start to latest end."""
    if isinstance(s, list):
        ref = s[0]
        return (max(ref.elapsed_time(x) for x in e) - min(ref.elapsed_time(x) for x in s))
    return s.elapsed_time(e)


def summary(window_s: float = None, units: int = None):
    """Per-kernel figures of the recorded launches.  The roofline object names
    the kernel (or library call) with the largest total time, whoever wrote
    it.  With the timed region's wall time ``window_s`` (and the number of
    updates ``units`` it holds) it also reports the headline efficiency
    sum_k max(F_k / P_k, B_k / BW) / window: the time every recorded launch
    would take at its bound's peak on its algorithmic work, over the time the
    region took."""
    if not _records:
        return None
    torch.cuda.synchronize()
    agg = defaultdict(lambda: [0, 0.0, 0.0, None, 0.0])
    for name, s, e, work, bound, nb in _records:
        a = agg[name]
        a[0] += 1
        a[1] += _elapsed_ms(s, e)
        a[2] += work
        a[3] = bound
        a[4] += nb
    kernels = {}
    ideal_s = 0.0
    for name, (cnt, ms, work, bound, nb) in agg.items():
        mean_ms = ms / cnt
        per = work / cnt
        if bound == "hbm":
            achieved = per / (mean_ms * 1e-3) / 1e9
            ideal = work / (PEAK["hbm"] * 1e9)
            ent = {"bound": bound, "achieved": achieved, "peak": PEAK[bound], "unit": UNIT[bound],
                   "frac": achieved / PEAK[bound]}
        else:
            t_flops = work / (PEAK[bound] * 1e12)
            t_bytes = nb / (PEAK["hbm"] * 1e9)
            ideal = max(t_flops, t_bytes)
            flops_rate = per / (mean_ms * 1e-3) / 1e12
            ent = {"bound": bound, "achieved": flops_rate, "peak": PEAK[bound],
                   "unit": UNIT[bound], "frac": flops_rate / PEAK[bound]}
            if t_bytes > t_flops:
                # the algorithm's bytes bound it: report against HBM
                bw = nb / cnt / (mean_ms * 1e-3) / 1e9
                ent = {"bound": "hbm", "achieved": bw, "peak": PEAK["hbm"], "unit": "GB/s",
                       "frac": bw / PEAK["hbm"], "bytes_per_launch": nb / cnt,
                       "flops_achieved": flops_rate, "flops_unit": UNIT[bound],
                       "flops_frac": flops_rate / PEAK[bound]}
        ideal_s += ideal
        kernels[name] = dict(launches=cnt, mean_ms=mean_ms, work_per_launch=per, total_ms=ms,
                             ideal_ms=ideal * 1e3, **ent)
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    out = {"kernel": dom, "bound": "hbm" if d["bound"] == "hbm" else "mfma",
           "achieved": d["achieved"], "peak": d["peak"], "unit": d["unit"],
           "frac": d["frac"], "traffic": None, "mean_launch_ms": d["mean_ms"],
           "all": kernels}
    if "bytes_per_launch" in d:
        out["bytes_per_launch"] = d["bytes_per_launch"]
        out["flops_frac"] = d["flops_frac"]
    if window_s:
        out["headline_efficiency"] = {
            "value": ideal_s / window_s,
            "what": "sum over the recorded launches of max(algorithmic flops / peak, "
                    "algorithmic bytes / 8 TB/s), divided by the timed region's wall time",
            "ideal_ms_per_update": ideal_s * 1e3 / units if units else None,
            "window_ms_per_update": window_s * 1e3 / units if units else None}
    return out
