"""Autotuner of the spherical engine's launch configuration.

The step's result does not depend on how it is launched (every schedule and
variant is bit-identical), so the tuner only times.  It searches, one
dimension at a time, the schedule (one kernel per sub-phase; one persistent
queue-driven kernel per RK stage with 1 or 2 workgroups per compute unit; the
sub-phase kernels with fluxes and hydro update fused per block, "slab"; the
sub-phase kernels with stage 1's con2prim fused into stage 0's update; or one
launch per stage with work-queue phases between grid barriers, with its
occupancy hint) and, for the sub-phase schedules, the XCD numbering, the workgroup size
and occupancy hint of the flux, update and con2prim kernels and the form of
the face-field update (one face per thread, or each edge's EMF once in LDS
with 2 to 16 workgroups per block); each candidate
is timed over a few graph-replayed steps of the caller's own runs.  The choice
is cached per (device, block shape, block count, precision) in a JSON file
(``$BHFLUID_TUNE_CACHE``, default ``~/.cache/bhfluid/tune.json``).
"""
from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np

from . import sph

DEFAULT = dict(schedule=0, wg_per_cu=0, flux=0, update=0, c2p=1, xcd=1, face=2)


def _cache_path() -> Path:
    p = os.environ.get("BHFLUID_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "bhfluid" / "tune.json"


def _key(eng: sph.Engine, runs, mixed: bool) -> str:
    from . import device_info
    nb = sum(r.nb for r in runs)
    prec = "mixed_fast" if mixed == 2 else "mixed" if mixed else "fp64"
    return f"{device_info().split(';')[0]}|bz{eng.bz}|nb{nb}|{prec}"


def apply(eng: sph.Engine, cfg: dict):
    eng.set_schedule(int(cfg.get("schedule", cfg.get("persistent", 0))), int(cfg["wg_per_cu"]))
    eng.set_variants(int(cfg["flux"]), int(cfg["update"]), int(cfg["c2p"]))
    eng.set_xcd_remap(bool(cfg.get("xcd", 1)))
    eng.set_face_variant(int(cfg.get("face", 2)))


def _time(eng, runs, cfg, steps: int) -> float:
    apply(eng, cfg)
    eng.step(runs, 2, want_dt=False)        # capture + warm
    ms = eng.step_timed(runs, steps)
    return float(np.median(ms))


def tune(eng: sph.Engine, runs: List[sph.Run], mixed: bool, steps: int = 4, use_cache: bool = True,
         log: Optional[list] = None) -> Dict:
    """Pick the fastest launch configuration for ``runs`` on ``eng`` (which
    must already hold their initial data; their state advances while the
    candidates are timed), apply it and return it with its timings."""
    key = _key(eng, runs, mixed)
    path = _cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        cfg = cache[key]["config"]
        apply(eng, cfg)
        return dict(config=cfg, cached=True, timings=cache[key]["timings"])
    timings = []
    best = dict(DEFAULT)
    best_ms = _time(eng, runs, best, steps)
    timings.append((dict(best), best_ms))
    for c in [dict(best, schedule=1, wg_per_cu=wg) for wg in (1, 2)] + [dict(best, schedule=s) for s in (2, 3, 4)]:
        t = _time(eng, runs, c, steps)
        timings.append((c, t))
        if t < best_ms:
            best, best_ms = c, t
    if best["schedule"] == 4:
        # the queue schedule takes its occupancy hint from the flux variant
        for v in (1, 3, 4):
            c = dict(best, flux=v)
            t = _time(eng, runs, c, steps)
            timings.append((c, t))
            if t < best_ms:
                best, best_ms = c, t
    elif best["schedule"] != 1:
        c = dict(best, xcd=0)
        t = _time(eng, runs, c, steps)
        timings.append((c, t))
        if t < best_ms:
            best, best_ms = c, t
        # the slab schedule has no separate update kernel and computes only
        # the side faces with k_flux (one face per thread)
        dims = ("flux", "c2p", "face") if best["schedule"] == 2 else ("flux", "update", "c2p", "face")
        for dim in dims:
            top = 12 if (dim == "flux" and best["schedule"] != 2) else 5 if dim == "face" else 6
            for v in range(0, top):
                if v == best[dim]:
                    continue
                c = dict(best, **{dim: v})
                t = _time(eng, runs, c, steps)
                timings.append((c, t))
                if t < best_ms:
                    best, best_ms = c, t
            if dim == "update" and best["schedule"] in (0, 3):
                # schedules 0 and 3 differ in the update kernel (3 runs
                # stage 1's con2prim inside stage 0's update), so each is
                # judged with its own best update variant
                other = 3 - best["schedule"]
                for v in range(0, 6):
                    c = dict(best, schedule=other, update=v)
                    t = _time(eng, runs, c, steps)
                    timings.append((c, t))
                    if t < best_ms:
                        best, best_ms = c, t
    apply(eng, best)
    out = dict(config=best, cached=False, timings=[dict(cfg=c, ms=t) for c, t in timings])
    if use_cache:
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            cache[key] = dict(config=best, timings=out["timings"])
            path.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    if log is not None:
        log.extend(out["timings"])
    return out
