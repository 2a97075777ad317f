"""Host loop over the stage-level entry points (``wtet_stage_*``).

``run`` is the control flow of the reference's ``phase1::run``
(src/phase1.rs:16-1431) written against the C ABI one dispatch group at a
time: vote (with the compaction of the uninserted list), split,
split_points, at most ``max_flip_iterations`` flip iterations, relocate;
then ``stage_finish`` does what ``delaunay_3d`` does after phase 1
(src/lib.rs:178-310).  ``wtet_delaunay`` runs the same kernels from one call
with one host read per round; this loop reads the counters after every
stage, which is what a foreign host loop or a unit test needs.
"""
from __future__ import annotations

from typing import Callable

from . import _core

MAX_ROUNDS = 4096


def run(ctx: "_core.Context", points, config: "_core.Config | None" = None,
        observe: Callable[[str, dict], None] | None = None) -> "_core.DelaunayResult":
    """Triangulates `points` through the stage entry points.

    observe(stage, info), when given, is called after every stage with the
    stage name and the counters after it (round, uninserted, winners,
    tet_slots, active, checked, flips, violations)."""
    cfg = config if config is not None else _core.Config()
    note = observe or (lambda stage, info: None)
    info = ctx.stage_init(points, cfg)
    note("init", info)
    try:
        while info["uninserted"] > 0:
            if info["round"] >= MAX_ROUNDS:
                raise RuntimeError("insertion did not finish")
            info = ctx.stage_vote()
            note("vote", info)
            info = ctx.stage_split()
            note("split", info)
            info = ctx.stage_split_points()
            note("split_points", info)
            if cfg.enable_flipping:
                for _ in range(max(int(cfg.max_flip_iterations), 1)):
                    if info["active"] == 0:
                        break
                    info = ctx.stage_flip()
                    note("flip", info)
                    if info["flips"] == 0:
                        break
                info = ctx.stage_relocate()
                note("relocate", info)
        return ctx.stage_finish()
    finally:
        ctx.stage_end()
