"""Launch-policy autotuner for the persistent orbit integrator.

A batch of orbits of uneven length runs as one persistent launch that is
bounded by its longest orbit's serial step chain, so how a caller keeps the
device busy matters as much as the kernel: how many independent batches are
in flight (each on its own stream), what fraction of the device each launch
takes, how many workgroups a compute unit holds, how many lanes start on the
head of the longest-first queue, and how many tickets a lane takes from the
queue at a time.  `tune_batches` searches these by coordinate descent
(vary one setting at a time from the current best, keep any change that is
faster by more than the noise margin, repeat until a sweep changes nothing),
timing short pipelined runs of the caller's own batch.  The result is cached
per device, library build and workload, so a process tunes once.  Batches in
flight beyond the process's hardware queues (GPU_MAX_HW_QUEUES, read when
HIP starts) share queues and run in turn, which the measurements show.

Orbits per lane is one by design (the state, stages and controller of an
orbit fill a lane's registers), and waves per SIMD is fixed at compile time
by the kernel's register budget; the tuner reports both.
"""

from __future__ import annotations

import ctypes
import hashlib
import json
import os
import time
from pathlib import Path

from . import _lib

# The search space.  blocks_per_cu above the kernel's occupancy is clipped by
# the library, so 1 and 2 (2 = the Milky Way kernel's occupancy at 256 lanes
# per workgroup) cover it.  Twelve batches in flight (three devices' worth of
# quarter-device launches) measured 19.0 ms per pass in one run and 23.7 in
# another on sixteen hardware queues, so the search stops at ten; launches
# over the whole device (parts 1) leave nothing for the next batch to fill
# and are left out, as is a single batch in flight.
SPACE = {
    "in_flight": (2, 3, 4, 6, 8, 10),
    "parts": (2, 3, 4, 6),
    "blocks_per_cu": (1, 2),
    "head_lanes": (0, 1024, 4096, 8192),
    "chunk": (1, 2, 4),
}
START = {"in_flight": 8, "parts": 4, "blocks_per_cu": 2, "head_lanes": 4096, "chunk": 1}
MAX_IN_FLIGHT = max(SPACE["in_flight"])


def apply(settings: dict) -> None:
    """Set the library's launch policy (the stream count is the caller's)."""
    L = _lib.lib()
    _lib.check(L.gx_set_launch_fraction(int(settings["parts"]), None), "gx_set_launch_fraction")
    _lib.check(L.gx_set_launch_shape(0, int(settings["blocks_per_cu"]), None, None), "gx_set_launch_shape")
    _lib.check(L.gx_set_queue_policy(int(settings["head_lanes"]), int(settings["chunk"]), None, None),
               "gx_set_queue_policy")


def current() -> dict:
    L = _lib.lib()
    parts = ctypes.c_int32()
    blk, bpc = ctypes.c_int32(), ctypes.c_int32()
    head, chunk = ctypes.c_int64(), ctypes.c_int32()
    _lib.check(L.gx_set_launch_fraction(0, ctypes.byref(parts)), "gx_set_launch_fraction")
    _lib.check(L.gx_set_launch_shape(0, 0, ctypes.byref(blk), ctypes.byref(bpc)), "gx_set_launch_shape")
    _lib.check(L.gx_set_queue_policy(-1, 0, ctypes.byref(head), ctypes.byref(chunk)), "gx_set_queue_policy")
    return {"parts": parts.value, "blocks_per_cu": bpc.value, "head_lanes": head.value, "chunk": chunk.value}


def _cache_path() -> Path:
    base = os.environ.get("GALAX_AMD_TUNE_CACHE")
    if base:
        return Path(base)
    return Path(os.environ.get("XDG_CACHE_HOME", Path.home() / ".cache")) / "galax_amd" / "tune.json"


def _build_id() -> str:
    st = _lib.LIB_PATH.stat()
    return hashlib.sha1(f"{st.st_size}:{int(st.st_mtime)}".encode()).hexdigest()[:12]


def device_key(torch) -> str:
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}|{p.multi_processor_count}cu|{_build_id()}"


def _load_cache() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(c: dict) -> None:
    path = _cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(c, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs a re-tune


def time_batches(torch, launch, settings: dict, passes: int, warmup: int, streams=None) -> float:
    """Wall milliseconds per pass of at least `passes` batches (and six per
    stream, so that the pipeline's fill and drain weigh little against its
    steady state) kept `in_flight` deep on as many streams,
    after at least `warmup` (and one per stream) untimed passes; launch(j)
    issues one pass on the current stream using buffer set j.  `streams`,
    when given, are the streams to use (the first in_flight of them): HIP
    maps each new stream onto one of the process's hardware queues in turn,
    so streams created per trial would end up sharing queues with one
    another and run their launches in turn; the caller creates its streams
    once and times with the same ones it will run on."""
    apply(settings)
    nfl = int(settings["in_flight"])
    passes = max(passes, 6 * nfl)
    warmup = max(warmup, nfl)
    if streams is None:
        streams = [torch.cuda.Stream() for _ in range(nfl)]
    streams = list(streams)[:nfl]
    if len(streams) < nfl:
        raise ValueError(f"{nfl} batches in flight need {nfl} streams, got {len(streams)}")
    main = torch.cuda.current_stream()
    for s in streams:
        s.wait_stream(main)
    for k in range(warmup):
        with torch.cuda.stream(streams[k % nfl]):
            launch(k % nfl)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for s in streams:
        s.wait_stream(main)
    for k in range(passes):
        with torch.cuda.stream(streams[k % nfl]):
            launch(k % nfl)
    for s in streams:
        main.wait_stream(s)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / passes


def coordinate_descent(measure, space: dict, start: dict, margin: float = 0.02, sweeps: int = 3):
    """Minimise measure(settings) over the product of `space`'s value lists:
    from `start`, try every other value of one setting at a time and move to
    it when it is faster by more than `margin` (relative); repeat until a
    sweep moves nothing (at most `sweeps`).  Each distinct setting is
    measured once.  Returns (best, best_value, trials)."""
    seen = {}
    trials = []

    def m(cfg):
        k = tuple(sorted(cfg.items()))
        if k not in seen:
            seen[k] = measure(cfg)
            trials.append({**cfg, "ms_per_pass": round(seen[k], 3)})
        return seen[k]

    best = dict(start)
    best_v = m(best)
    for _ in range(sweeps):
        changed = False
        for name, values in space.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best, **{name: v})
                val = m(cand)
                if val < best_v * (1.0 - margin):
                    best, best_v, changed = cand, val, True
        if not changed:
            break
    return best, best_v, trials


def confirm_winner(measure, start: dict, best: dict, best_v: float, margin: float, rounds: int = 2):
    """A pass lasts as long as its longest orbit's chain, so single trials
    are noisy: a setting that won the search must beat the start again in
    `rounds` alternating re-measurements (means compared, with the search's
    margin) before it replaces it.  Returns (settings, value, record)."""
    if best == start:
        return best, best_v, None
    tb, ts = [], []
    for _ in range(rounds):
        ts.append(measure(start))
        tb.append(measure(best))
        # (alternating, so that drift in the device's state hits both)
    mb, ms = sum(tb) / len(tb), sum(ts) / len(ts)
    keep = mb < ms * (1.0 - margin)
    chosen, value = (dict(best), mb) if keep else (dict(start), ms)
    record = {"start_ms": [round(x, 3) for x in ts], "best_ms": [round(x, 3) for x in tb], "kept": chosen}
    return chosen, value, record


def tune_batches(torch, launch, workload_key: str, *, max_in_flight: int, passes: int = 12, warmup: int = 3,
                 margin: float = 0.02, use_cache: bool = True, streams=None, log=None) -> dict:
    """Coordinate-descent search of the launch policy for repeated passes of
    one batch.  `launch(j)` issues one pass with buffer set j (0 <=
    j < max_in_flight); `streams` (max_in_flight of them, created once by
    the caller) are the streams the trials run on.  Returns {"settings", "ms_per_pass", "trials",
    "confirmation", "seconds", "cached", "search", "fixed"}; the best
    settings are left applied."""
    key = f"{device_key(torch)}|{workload_key}"
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        res = dict(cache[key])
        res["cached"] = True
        apply(res["settings"])
        return res
    t_start = time.perf_counter()
    space = dict(SPACE)
    space["in_flight"] = tuple(v for v in SPACE["in_flight"] if v <= max_in_flight)
    start = dict(START)
    start["in_flight"] = min(start["in_flight"], max_in_flight)

    def measure(cfg):
        ms = time_batches(torch, launch, cfg, passes, warmup, streams)
        if log:
            log(f"[tune] {cfg} -> {ms:.2f} ms/pass")
        return ms

    best, best_ms, trials = coordinate_descent(measure, space, start, margin)
    best, best_ms, confirm = confirm_winner(measure, start, best, best_ms, margin)
    apply(best)
    res = {
        "settings": best,
        "ms_per_pass": best_ms,
        "trials": trials,
        "confirmation": confirm,
        "seconds": time.perf_counter() - t_start,
        "cached": False,
        "search": "coordinate descent over in_flight x parts x blocks_per_cu x head_lanes x chunk from "
                  f"{start}, max({passes}, 6 x in_flight) timed passes per trial, keep a change only if > "
                  f"{margin:.0%} faster, then the winner against the start twice more, alternating",
        "fixed": {"orbits_per_lane": 1, "waves_per_simd": "compile-time (register budget)"},
    }
    if use_cache:
        cache[key] = {k: res[k] for k in ("settings", "ms_per_pass", "trials", "confirmation", "seconds", "search",
                                          "fixed")}
        _store_cache(cache)
    return res
