"""Tile autotuner of the grouped launch: vectors per workgroup and layers per LDS stage.

The grouped kernel takes a tile per mesh: ``w`` waveguides per lane (fixed by N), ``v`` vectors
per lane (vectors per workgroup = waves x lane groups x v) and ``chunk`` layers per LDS
reduction stage of the adjoint.  ``tune_group`` sweeps (v, chunk) per width class W by
coordinate descent over the classes (largest total work first), timing the whole group with HIP
events on each candidate, and returns the best setting per class together with the sweep log.
Candidates are the tiles the grouped kernel compiles: v in {1, 2, 4} for W = 2, {1, 2} for
W = 4, {1} for W = 8 and 16; chunk a multiple of the adjoint's unrolled group (4 layers for
W <= 4, 2 otherwise) that fits the 64 KiB stage.
"""
from __future__ import annotations

import json
import os
from typing import Dict, List, Tuple

import numpy as np
import torch

from .batched import GroupJob, MeshGroup

V_CHOICES = {2: (1, 2, 4), 4: (1, 2), 8: (1,), 16: (1,)}
STAGE_BYTES = 64 * 1024


def width_class(N: int) -> int:
    return 2 if N <= 128 else 4 if N <= 256 else 8 if N <= 512 else 16


def _chunks(W: int, N: int) -> List[int]:
    waves = 2 if W == 16 else 4
    lpv = 1
    while lpv < (N + W - 1) // W:
        lpv *= 2
    per_layer = waves * lpv * W * 16
    depth = 4 if W <= 4 else 2
    cmax = min(16, STAGE_BYTES // per_layer)
    out, c = [], depth
    while c <= cmax:
        out.append(c)
        c *= 2
    return out or [depth]


def _apply(jobs: List[GroupJob], cfg: Dict[int, Tuple[int, int]]):
    for j in jobs:
        W = width_class(j.mesh.N)
        v, c = cfg.get(W, (0, 0))
        j.tile = (W, v, 0, c)


def _time(group: MeshGroup, reps: int) -> float:
    group.run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        group.run()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


def tune_group(jobs: List[GroupJob], reps: int = 5, log=None):
    """Pick (v, chunk) per width class for ``jobs`` (tiles are written into the jobs)."""
    work = {}
    for j in jobs:
        W = width_class(j.mesh.N)
        work[W] = work.get(W, 0.0) + float(j.B) * j.mesh.L * j.mesh.N
    classes = sorted(work, key=lambda w: -work[w])
    cfg: Dict[int, Tuple[int, int]] = {W: (0, 0) for W in classes}
    _apply(jobs, cfg)
    grp = MeshGroup(jobs)
    best = _time(grp, reps)
    sweep = [{"cfg": {str(k): v for k, v in cfg.items()}, "ms": best}]
    for W in classes:
        Ns = sorted({j.mesh.N for j in jobs if width_class(j.mesh.N) == W})
        chunks = sorted(set(c for N in Ns for c in _chunks(W, N)))
        for v in V_CHOICES[W]:
            for c in chunks:
                trial = dict(cfg)
                trial[W] = (v, c)
                _apply(jobs, trial)
                grp.replan(jobs)
                t = _time(grp, reps)
                sweep.append({"cfg": {str(k): vv for k, vv in trial.items()}, "ms": t})
                if t < best:
                    best, cfg = t, trial
    _apply(jobs, cfg)
    grp.close()
    if log:
        with open(log, "w") as f:
            json.dump({"best": {str(k): v for k, v in cfg.items()}, "best_ms": best, "sweep": sweep}, f, indent=1)
    return cfg, best, sweep


def load_or_tune(jobs: List[GroupJob], cache: str, reps: int = 5):
    """Tile settings from a JSON cache written by an earlier sweep, else sweep and store."""
    if cache and os.path.exists(cache):
        with open(cache) as f:
            cfg = {int(k): tuple(v) for k, v in json.load(f)["best"].items()}
        _apply(jobs, cfg)
        return cfg, None, []
    return tune_group(jobs, reps=reps, log=cache)
