"""Tile-shape autotuner for the fused stepping kernels.

For an (equation, integrator, precision, grid-size class) key it times every
kernel variant ``libpml.so`` offers for that combination (tile height, fused
steps per launch, workgroup size -- ``pml_fdm_configs`` /
``pml_fdm_config_name``) on the live device, on a scratch batch of the
caller's shape, and keeps the fastest.  Results persist as JSON in
``$PML_AUTOTUNE_CACHE`` (default ``~/.cache/pararealml_amd/autotune.json``)
and are reused by ``FDMOperator`` when it is not given a ``tile_config``.
The cache stores the variant's *name*; a lookup resolves it against the
variants the loaded library offers, so an entry written by another build
(pruned or reordered variants) falls back to the default variant instead
of selecting a wrong or out-of-range one.
"""

import json
import os
import threading
from pathlib import Path
from typing import Dict, List, Optional, Sequence

_lock = threading.Lock()
_cache: Optional[Dict[str, str]] = None


def _cache_path() -> Path:
    env = os.environ.get("PML_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "pararealml_amd" / "autotune.json"


def _load() -> Dict[str, str]:
    global _cache
    if _cache is None:
        path = _cache_path()
        try:
            raw = json.loads(path.read_text())
            # entries that are not variant names (e.g. bare indices) are stale
            _cache = {k: v for k, v in raw.items() if isinstance(v, str)}
        except (OSError, ValueError, AttributeError):
            _cache = {}
    return _cache


def variant_names(eq_id: int, integrator: int, prec: int) -> List[str]:
    """The kernel variants the loaded library offers, in index order."""
    from pararealml_amd import _native as nat

    lib = nat.load_library()
    return [
        lib.pml_fdm_config_name(eq_id, integrator, prec, i).decode()
        for i in range(lib.pml_fdm_configs(eq_id, integrator, prec))
    ]


def _size_class(cells: int) -> str:
    for bound, name in ((1 << 14, "xs"), (1 << 17, "s"), (1 << 20, "m"), (1 << 23, "l")):
        if cells <= bound:
            return name
    return "xl"


def key_for(problem, integrator: int, precision: str, cells: Optional[int] = None) -> str:
    cells = problem.plane if cells is None else cells
    return f"{problem.eq_name}/{integrator}/{precision}/{problem.ndim}d/{_size_class(cells)}"


def lookup(problem, integrator: int, precision: str, cells: Optional[int] = None,
           names: Optional[Sequence[str]] = None) -> int:
    """Index of the tuned variant for this key among ``names`` (default: the
    loaded library's variants), or 0 (the default variant) when nothing is
    recorded or the recorded name is not offered."""
    with _lock:
        name = _load().get(key_for(problem, integrator, precision, cells))
    if name is None:
        return 0
    if names is None:
        from pararealml_amd import _device as dev
        from pararealml_amd import _native as nat

        names = variant_names(nat.PML_EQ[problem.eq_name], integrator,
                              dev.precision_id(precision))
    try:
        return list(names).index(name)
    except ValueError:
        return 0


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
        best_name = lib.pml_fdm_config_name(batch.eq_id, batch.integrator, batch.prec,
                                            best).decode()
        with _lock:
            data = _load()
            data[key] = best_name
            path = _cache_path()
            try:
                path.parent.mkdir(parents=True, exist_ok=True)
                path.write_text(json.dumps(data, indent=1, sort_keys=True))
            except OSError:
                pass
    return best, timings
