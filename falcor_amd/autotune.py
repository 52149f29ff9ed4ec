"""Launch-configuration autotuner for the traversal kernels.

The BVH4 traversal kernels are compiled for several (threads per block, LDS stack entries per
lane) pairs and run on a persistent grid of `traceBlocksPerCU` blocks per CU. Which combination
is fastest depends on the scene (stack depth, ray coherence) and on the frame shape (rays per
launch), so a pass is tuned on its own workload: `tune` renders a few frames per candidate with
the pass's own properties and keeps the fastest, timing with HIP events on the pass's stream.

Tuned results can be kept in a JSON file (`save`/`load`) keyed by GPU architecture, pass kind
and a workload tag chosen by the caller.
"""
from __future__ import annotations

import json
import time
from pathlib import Path

LAUNCH_KEYS = ("traceBlockSize", "traceStackDepth", "traceBlocksPerCU")
BLOCK_SIZES = (64, 128, 256)
STACK_DEPTHS = (16, 24, 32)
THREADS_PER_CU = (512, 1024, 1536, 2048)
DEFAULT_FILE = Path(__file__).resolve().parent / "tuned_launch.json"


def candidates(block_sizes=BLOCK_SIZES, stack_depths=STACK_DEPTHS, threads_per_cu=THREADS_PER_CU):
    """All (block, lds, blocksPerCU) triples the library accepts from the given axes."""
    out = []
    for b in block_sizes:
        for d in stack_depths:
            for t in threads_per_cu:
                if t % b == 0 and t // b >= 1:
                    out.append({"traceBlockSize": b, "traceStackDepth": d, "traceBlocksPerCU": t // b})
    return out


def _time_frames(run, reps: int) -> float:
    import torch

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for i in range(reps):
        run(i)
    ev1.record()
    ev1.synchronize()
    return ev0.elapsed_time(ev1) / reps


def tune(pass_, run, configs=None, reps: int = 3, warmup: int = 1, log=None) -> dict:
    """Sets each candidate launch configuration on `pass_`, calls `run(i)` (one frame on the
    current stream) `warmup` + `reps` times and leaves the fastest configuration set.

    Returns {"best": {key: value}, "ms": best ms per frame, "table": [(config, ms), ...]}.
    """
    configs = candidates() if configs is None else configs
    saved = {k: pass_.get_property(k) for k in LAUNCH_KEYS if k in configs[0]} if configs else {}
    table = []
    for cfg in configs:
        apply(pass_, cfg)
        for i in range(warmup):
            run(i)
        ms = _time_frames(run, reps)
        table.append((dict(cfg), ms))
        if log:
            log(f"autotune {cfg}: {ms:.3f} ms")
    if not table:
        apply(pass_, saved)
        raise ValueError("no launch configurations to try")
    best, best_ms = min(table, key=lambda r: r[1])
    apply(pass_, best)
    return {"best": best, "ms": best_ms, "table": table}


def _arch() -> str:
    import torch

    p = torch.cuda.get_device_properties(0)
    return getattr(p, "gcnArchName", p.name).split(":")[0]


def save(kind: str, tag: str, best: dict, path: Path | str = DEFAULT_FILE, arch: str | None = None):
    path = Path(path)
    db = json.loads(path.read_text()) if path.exists() else {}
    db.setdefault(arch or _arch(), {}).setdefault(kind, {})[tag] = {k: int(best[k]) for k in LAUNCH_KEYS if k in best} | {
        "tuned": time.strftime("%Y-%m-%d")}
    path.write_text(json.dumps(db, indent=1, sort_keys=True) + "\n")


def load(kind: str, tag: str, path: Path | str = DEFAULT_FILE, arch: str | None = None) -> dict | None:
    path = Path(path)
    if not path.exists():
        return None
    db = json.loads(path.read_text())
    ent = db.get(arch or _arch(), {}).get(kind, {}).get(tag)
    return None if ent is None else {k: int(ent[k]) for k in LAUNCH_KEYS if k in ent}


def apply(pass_, cfg: dict):
    """Sets a launch configuration returned by `load` or `tune` (blocks per CU first lowered to
    1 so that every intermediate state satisfies blocksPerCU * blockSize <= 2048)."""
    pass_.set_property("traceBlocksPerCU", 1)
    pass_.set_property("traceBlockSize", cfg["traceBlockSize"])
    pass_.set_property("traceStackDepth", cfg["traceStackDepth"])
    pass_.set_property("traceBlocksPerCU", cfg["traceBlocksPerCU"])
