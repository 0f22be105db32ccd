"""Autotuner for the point traversal: picks the BVH leaf size, the ray reordering and the
lane-refill run length (rays per 64-lane wave) that minimise the measured traversal time
on a sample of the query's rays.

Every candidate gives identical results (the traversal collects all hits whatever the
tree shape, reordering only permutes rays, and a refilled lane runs the same per-ray
program), so the tuner changes speed only. The stack depth follows from the tree height
(the kernels are instantiated for 24/32/48/64 entries) and a workgroup is one wave."""
from __future__ import annotations

import torch

from . import _lib
from .device import Bvh, trace_points


def tune_traversal(ctx, vertices, indices, tri_to_obj, sample_points, axis: int = 2,
                   leaf_sizes=(1, 2, 4, 8), rays_per_wave=(0, 64, 128, 256), reps: int = 3):
    """Times every (leaf size, reordering on/off, rays per wave) on `sample_points` with HIP
    events on the context's stream. Returns (best trial, all trials); each trial is a dict
    {"leaf_size", "ray_sort", "rays_per_wave", "ms", "depth"}. The context's options are
    restored afterwards; apply the winner with apply()."""
    trials = []
    n = sample_points.shape[0]
    out = torch.empty(n, dtype=torch.int32, device=sample_points.device)
    saved = (ctx.get_option(_lib.OPT_RAY_SORT_MIN), ctx.get_option(_lib.OPT_RAYS_PER_WAVE))
    try:
        for leaf in leaf_sizes:
            bvh = Bvh(ctx, vertices, indices, tri_to_obj, leaf_size=leaf)
            depth = int(bvh.info().max_depth)
            for sort in (True, False):
                ctx.set_option(_lib.OPT_RAY_SORT_MIN, 1 if sort else 0)
                for pw in rays_per_wave:
                    ctx.set_option(_lib.OPT_RAYS_PER_WAVE, pw)
                    trace_points(ctx, bvh, sample_points, axis=axis, out=out)  # warm
                    ctx.profile(True)
                    ctx.profile_read()
                    for _ in range(reps):
                        trace_points(ctx, bvh, sample_points, axis=axis, out=out, sync=False)
                    ctx.sync()
                    ms, k = ctx.profile_read()
                    ctx.profile(False)
                    trials.append({"leaf_size": leaf, "ray_sort": sort, "rays_per_wave": pw,
                                   "ms": round(ms / max(k, 1), 4), "depth": depth})
            bvh.free()
    finally:
        ctx.set_option(_lib.OPT_RAY_SORT_MIN, saved[0])
        ctx.set_option(_lib.OPT_RAYS_PER_WAVE, saved[1])
    best = min(trials, key=lambda t: t["ms"])
    return best, trials


def apply(ctx, trial) -> None:
    """Sets the context options of a trial (the leaf size is a BVH build parameter)."""
    ctx.set_option(_lib.OPT_RAY_SORT_MIN, 1 if trial["ray_sort"] else 0)
    ctx.set_option(_lib.OPT_RAYS_PER_WAVE, trial["rays_per_wave"])
