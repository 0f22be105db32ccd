"""Autotuner for the point traversal: picks between the axis grid and the BVH walk, and for
the walk the BVH leaf size, the ray reordering and the lane-refill run length (rays per
64-lane wave), for the grid its run length -- whichever minimises the measured time of a
point launch on a sample of the query's rays.

Every candidate gives identical results (the walk collects all hits whatever the tree
shape, the grid tests exactly the triangles whose leaf box the walk accepts, reordering only
permutes rays, and a refilled lane runs the same per-ray program), so the tuner changes
speed only. Each trial is timed over whole launches (reordering included) on the host clock
after a warm-up launch; the grid and the trees are built before timing starts."""
from __future__ import annotations

import time

import torch

from . import _lib
from .device import Bvh, axis_grid_prepare, trace_points

_OPTS = (_lib.OPT_RAY_SORT_MIN, _lib.OPT_RAYS_PER_WAVE, _lib.OPT_AXIS_GRID, _lib.OPT_GRID_RAYS_PER_WAVE)


def _time(ctx, bvh, pts, axis, out, reps):
    trace_points(ctx, bvh, pts, axis=axis, out=out)  # warm (and builds the grid if allowed)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        trace_points(ctx, bvh, pts, axis=axis, out=out, sync=False)
    ctx.sync()
    return (time.perf_counter() - t0) * 1e3 / reps


def tune_traversal(ctx, vertices, indices, tri_to_obj, sample_points, axis: int = 2,
                   leaf_sizes=(1, 2, 4, 8), rays_per_wave=(0, 64, 128, 256), reps: int = 3,
                   grid_rays_per_wave=(256, 1024, 4096)):
    """Times the axis grid (each grid run length, when a grid is accepted) and the BVH walk
    (every leaf size x reordering on/off x rays per wave) on `sample_points`. Returns (best
    trial, all trials); a trial is {"axis_grid", "leaf_size", "ray_sort", "rays_per_wave",
    "grid_rays_per_wave", "ms", "depth"}. The context's options are restored afterwards;
    apply the winner with apply()."""
    trials = []
    n = sample_points.shape[0]
    out = torch.empty(n, dtype=torch.int32, device=sample_points.device)
    saved = tuple(ctx.get_option(o) for o in _OPTS)
    try:
        for li, leaf in enumerate(leaf_sizes):
            bvh = Bvh(ctx, vertices, indices, tri_to_obj, leaf_size=leaf)
            depth = int(bvh.info().max_depth)
            if li == 0:
                info = axis_grid_prepare(ctx, bvh, axis)
                if info["state"] == 1:
                    ctx.set_option(_lib.OPT_AXIS_GRID, 1)
                    for gpw in grid_rays_per_wave:
                        ctx.set_option(_lib.OPT_GRID_RAYS_PER_WAVE, gpw)
                        trials.append({"axis_grid": True, "leaf_size": leaf, "ray_sort": False,
                                       "rays_per_wave": saved[1], "grid_rays_per_wave": gpw,
                                       "ms": round(_time(ctx, bvh, sample_points, axis, out, reps), 4),
                                       "depth": depth})
            ctx.set_option(_lib.OPT_AXIS_GRID, 0)
            for sort in (True, False):
                ctx.set_option(_lib.OPT_RAY_SORT_MIN, 1 if sort else 0)
                for pw in rays_per_wave:
                    ctx.set_option(_lib.OPT_RAYS_PER_WAVE, pw)
                    trials.append({"axis_grid": False, "leaf_size": leaf, "ray_sort": sort, "rays_per_wave": pw,
                                   "grid_rays_per_wave": saved[3],
                                   "ms": round(_time(ctx, bvh, sample_points, axis, out, reps), 4),
                                   "depth": depth})
            bvh.free()
    finally:
        for o, v in zip(_OPTS, saved):
            ctx.set_option(o, v)
    best = min(trials, key=lambda t: t["ms"])
    return best, trials


def apply(ctx, trial) -> None:
    """Sets the context options of a trial (the leaf size is a BVH build parameter)."""
    ctx.set_option(_lib.OPT_AXIS_GRID, 1 if trial["axis_grid"] else 0)
    ctx.set_option(_lib.OPT_RAY_SORT_MIN, 1 if trial["ray_sort"] else 0)
    ctx.set_option(_lib.OPT_RAYS_PER_WAVE, trial["rays_per_wave"])
    ctx.set_option(_lib.OPT_GRID_RAYS_PER_WAVE, trial["grid_rays_per_wave"])
