"""Launch-shape autotuner: classify workgroup size x instances per thread, and
the Hi-Z tile edge, measured on the GPU and cached as JSON per device.

The cache is keyed by the device's CU count and the workload signature; the
batched launch, the plain pass and the server read their shapes from the
context, which ``apply`` sets.
"""
from __future__ import annotations

import ctypes as C
import json
import time
from pathlib import Path

from ._lib import TuneResult, call
from .passes import BatchCull, Context, HiZBuildPass, Stream

CACHE = Path(__file__).resolve().parent / "tuned" / "shapes.json"


def get_shapes(ctx: Context) -> dict:
    wg, ipt, tile = C.c_uint32(), C.c_uint32(), C.c_uint32()
    call("hv_ctx_get_shapes", ctx.handle, C.byref(wg), C.byref(ipt), C.byref(tile))
    return {"wg": wg.value, "ipt": ipt.value, "hiz_tile": tile.value}


def apply(ctx: Context, shapes: dict) -> None:
    call("hv_ctx_set_shapes", ctx.handle, int(shapes["wg"]), int(shapes["ipt"]), int(shapes["hiz_tile"]))


def tune_batch(batch: BatchCull, reps: int = 5, replay: bool = False) -> dict:
    """Times every launch shape on `batch`: eager launches (the three kernels'
    sum), or with `replay` replays of its captured graph (as a renderer runs
    it every frame)."""
    res = (TuneResult * 16)()
    n, best = C.c_uint32(), TuneResult()
    call("hv_autotune_batch_replay" if replay else "hv_autotune_batch", batch.handle, reps, res, 16, C.byref(n),
         C.byref(best))
    return {"best": {"wg": best.wg, "ipt": best.ipt, "classify_ms": best.classify_ms, "total_ms": best.total_ms},
            "table": [{"wg": r.wg, "ipt": r.ipt, "classify_ms": r.classify_ms, "total_ms": r.total_ms}
                      for r in res[: n.value]]}


def tune_hiz(hiz: HiZBuildPass, reps: int = 10) -> dict:
    ms = (C.c_float * 2)()
    tile = C.c_uint32()
    call("hv_autotune_hiz", hiz.handle, reps, ms, C.byref(tile))
    return {"best_tile": tile.value, "ms_tile32": ms[0], "ms_tile64": ms[1]}


def key(cu_count: int, workload: str) -> str:
    """Cache key: the device's compute-unit count and the workload signature."""
    return f"cu{int(cu_count)}:{workload}"


def load_cache(path: Path = CACHE) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def lookup(cu_count: int, workload: str, path: Path = CACHE) -> dict | None:
    return load_cache(path).get(key(cu_count, workload))


def store(cu_count: int, workload: str, entry: dict, path: Path = CACHE) -> None:
    data = load_cache(path)
    entry = dict(entry, tuned_at=time.strftime("%Y-%m-%dT%H:%M:%S"))
    data[key(cu_count, workload)] = entry
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".json.tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    tmp.replace(path)


def cull_shape(ctx: Context, workload: str, prepare, reps: int = 5, path: Path = CACHE,
               retune: bool = False, replay: bool = False) -> tuple[dict, "BatchCull", bool]:
    """The classify launch shape for `workload` on this device, applied to the
    context, and a batch prepared with it. `prepare(ctx)` builds and prepares
    the batch the shape is tuned on. A cached entry is applied before the
    batch is built; on a miss (or retune) every shape is timed on the batch
    (hv_autotune_batch, which leaves the batch and the context on the
    fastest) and the result is stored; with `replay` the shapes are timed as
    graph replays (the cache keeps the two kinds apart). Returns (shape,
    batch, from_cache)."""
    cu, _ = ctx.device_info()
    if replay:
        workload = f"{workload}:replay"
    hit = None if retune else lookup(cu, workload, path)
    if hit is not None:
        apply(ctx, dict(get_shapes(ctx), wg=hit["wg"], ipt=hit["ipt"]))
        return hit, prepare(ctx), True
    batch = prepare(ctx)
    res = tune_batch(batch, reps, replay)
    store(cu, workload, dict(res["best"], table=res["table"]), path)
    return res["best"], batch, False


def steady_step_ms(ctx: Context, prepare, in_flight: int, steps: int, warmup: int = 4) -> float:
    """Milliseconds per step with `in_flight` prepared copies of the batch
    replaying their graphs on their own streams, steps taking turns (one
    step's column scan and scatter beside the next steps' classify)."""
    copies = [prepare(ctx) for _ in range(in_flight)]
    streams = [Stream(ctx) for _ in copies]
    try:
        for i in range(max(warmup, in_flight)):
            copies[i % in_flight].execute(stream=streams[i % in_flight].handle, graph=True)
        for b in copies:
            b.wait()
        t0 = time.perf_counter()
        for i in range(steps):
            copies[i % in_flight].execute(stream=streams[i % in_flight].handle, graph=True)
        for b in copies:
            b.wait()
        return (time.perf_counter() - t0) / steps * 1e3
    finally:
        for b in copies:
            b.close()
        for st in streams:
            st.close()


def cull_shape_in_flight(ctx: Context, workload: str, prepare, in_flight: int = 2, candidates: int = 3,
                         steps: int = 60, reps: int = 5, path: Path = CACHE,
                         retune: bool = False) -> tuple[dict, "BatchCull", bool]:
    """The classify launch shape for batches replayed `in_flight` at a time
    (steady_step_ms): the `candidates` shapes with the smallest eager
    three-kernel sum (cull_shape's table) are each timed in that mode and the
    fastest is applied, cached under its own key. Returns (shape, batch,
    from_cache); the entry holds every candidate's step time."""
    cu, _ = ctx.device_info()
    wl = f"{workload}:in_flight{in_flight}"
    hit = None if retune else lookup(cu, wl, path)
    if hit is not None:
        apply(ctx, dict(get_shapes(ctx), wg=hit["wg"], ipt=hit["ipt"]))
        return hit, prepare(ctx), True
    eager, batch, _ = cull_shape(ctx, workload, prepare, reps, path, retune)
    batch.close()
    table = (lookup(cu, workload, path) or {}).get("table") or [eager]
    cands = sorted(table, key=lambda r: r["total_ms"])[:max(1, candidates)]
    timed = []
    for c in cands:
        apply(ctx, dict(get_shapes(ctx), wg=c["wg"], ipt=c["ipt"]))
        timed.append({"wg": c["wg"], "ipt": c["ipt"], "total_ms": c["total_ms"],
                      "step_ms": steady_step_ms(ctx, prepare, in_flight, steps)})
    best = min(timed, key=lambda r: r["step_ms"])
    apply(ctx, dict(get_shapes(ctx), wg=best["wg"], ipt=best["ipt"]))
    entry = dict(best, in_flight=in_flight, candidates=timed)
    store(cu, wl, entry, path)
    return entry, prepare(ctx), False


def hiz_tile(ctx: Context, workload: str, hiz: HiZBuildPass, reps: int = 10, path: Path = CACHE,
             retune: bool = False) -> tuple[dict, bool]:
    """The Hi-Z tile edge for `workload` (the pyramid's size and mode) on this
    device, applied to the context: cached, or timed on `hiz` (which must have
    depth bound) and stored. Returns (entry, from_cache)."""
    cu, _ = ctx.device_info()
    hit = None if retune else lookup(cu, workload, path)
    if hit is not None:
        apply(ctx, dict(get_shapes(ctx), hiz_tile=hit["best_tile"]))
        return hit, True
    res = tune_hiz(hiz, reps)
    store(cu, workload, res, path)
    return res, False
