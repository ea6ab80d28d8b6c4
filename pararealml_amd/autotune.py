"""Tile-shape autotuner for the fused stepping kernels.

For an (equation, integrator, precision, grid-size class) key it times every
kernel variant ``libpml.so`` offers for that combination (tile height, fused
steps per launch, workgroup size -- ``pml_fdm_configs`` /
``pml_fdm_config_name``) on the live device, on a scratch batch of the
caller's shape, and keeps the fastest.  Results persist as JSON in
``$PML_AUTOTUNE_CACHE`` (default ``~/.cache/pararealml_amd/autotune.json``)
and are reused by ``FDMOperator`` when it is not given a ``tile_config``.
"""

import json
import os
import threading
from pathlib import Path
from typing import Dict, Optional

_lock = threading.Lock()
_cache: Optional[Dict[str, int]] = None


def _cache_path() -> Path:
    env = os.environ.get("PML_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "pararealml_amd" / "autotune.json"


def _load() -> Dict[str, int]:
    global _cache
    if _cache is None:
        path = _cache_path()
        try:
            _cache = {k: int(v) for k, v in json.loads(path.read_text()).items()}
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _size_class(cells: int) -> str:
    for bound, name in ((1 << 14, "xs"), (1 << 17, "s"), (1 << 20, "m"), (1 << 23, "l")):
        if cells <= bound:
            return name
    return "xl"


def key_for(problem, integrator: int, precision: str, cells: Optional[int] = None) -> str:
    cells = problem.plane if cells is None else cells
    return f"{problem.eq_name}/{integrator}/{precision}/{problem.ndim}d/{_size_class(cells)}"


def lookup(problem, integrator: int, precision: str, cells: Optional[int] = None) -> int:
    """The tuned variant for this key, or 0 (the default variant)."""
    with _lock:
        return _load().get(key_for(problem, integrator, precision, cells), 0)


def tune(batch, d_t: float, n_steps: int = 64, repeats: int = 3, store: bool = True):
    """Times every variant on ``batch`` (an ``FDMBatch``; its state is
    advanced and therefore clobbered) and records the fastest.  Returns
    ``(best_config, {config_name: seconds_per_step})``."""
    import torch

    lib = batch.lib
    n_cfg = lib.pml_fdm_configs(batch.eq_id, batch.integrator, batch.prec)
    timings = {}
    best, best_t = 0, float("inf")
    for cfg in range(n_cfg):
        name = lib.pml_fdm_config_name(batch.eq_id, batch.integrator, batch.prec, cfg).decode()
        batch.advance(d_t, 8, tile_config=cfg)  # warm-up
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        ts = []
        for _ in range(repeats):
            start.record(batch.stream)
            batch.advance(d_t, n_steps, tile_config=cfg, use_graph=True)
            end.record(batch.stream)
            end.synchronize()
            ts.append(start.elapsed_time(end) / 1e3 / n_steps)
        t = min(ts)
        timings[name] = t
        if t < best_t:
            best, best_t = cfg, t
    if store:
        cells = sum(batch.problems[i].plane for i in batch.grid_problem)
        key = key_for(batch.problems[0], batch.integrator, batch.precision, cells)
        with _lock:
            data = _load()
            data[key] = best
            path = _cache_path()
            try:
                path.parent.mkdir(parents=True, exist_ok=True)
                path.write_text(json.dumps(data, indent=1, sort_keys=True))
            except OSError:
                pass
    return best, timings
