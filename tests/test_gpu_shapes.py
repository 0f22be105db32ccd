"""The batched and served forms at the shapes of their workloads, checked against the
oracle (not only batched against per-query):

* batched: 256 independent queries (3/4 containment, 1/4 overlap), mesh sizes log-uniform
  in [1e3, 5e5] triangles and point counts log-uniform in [1e3, 1e7] (about 258 M points),
  run eagerly and as a captured graph; every query equals its own per-query launch and
  the oracle, every point and every pair;
* served: 64 resident datasets of the many-spheres generator (10,000 level-2 spheres,
  3.2 M triangles each), grid histograms at resolution 128, a seeded Poisson stream of
  containment, overlap and intersection requests (request meshes of 1-64 anisotropic,
  rotated level 2-4 icospheres), served with continuous batching under a memory budget
  that lets only some datasets get an axis grid; every request equals the same request
  served alone and the oracle;
* a server whose budget is smaller than one axis grid builds none and still answers
  every request like the oracle.
"""
import os

import numpy as np
import pytest
import torch

import oracle as orc
from rayspace3d_amd import _lib, gen
from rayspace3d_amd.device import (Bvh, CapturedBatch, DevEdges, Server, batch_execute, overlap_join,
                                   to_dev, to_host, trace_points)

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, os.cpu_count() or 1))


# ---- batched form: 256 queries at full size ----------------------------------------------

@pytest.mark.timeout(900)
def test_batched_workload_full_size_equals_per_query_and_oracle(ctx):
    print("batched workload: generating 256 queries", flush=True)
    qs = gen.batched_workload(seed=4, n_queries=256)
    queries, keep = [], []
    for q in qs:
        if q["kind"] == "contain":
            bvh = Bvh.from_mesh(ctx, q["mesh"])
            dp = to_dev(q["points"])
            queries.append({"kind": "contain", "bvh": bvh, "points": dp, "axis": 2,
                            "out": torch.full((dp.shape[0],), -7, dtype=torch.int32, device="cuda")})
            keep.append(bvh)
        else:
            ba, bb = Bvh.from_mesh(ctx, q["a"]), Bvh.from_mesh(ctx, q["b"])
            ea, eb = DevEdges(q["a"].edges).spatial(ctx), DevEdges(q["b"].edges).spatial(ctx)
            queries.append({"kind": "overlap", "bvh1": ba, "edges1": ea, "bvh2": bb, "edges2": eb})
            keep += [ba, bb]
    n_pts = sum(q["points"].shape[0] for q in queries if q["kind"] == "contain")
    print(f"batched workload: {len(queries)} queries, {n_pts} points; eager batch", flush=True)
    eager_pairs = batch_execute(ctx, queries)
    eager_ids = [to_host(q["out"]).copy() for q in queries if q["kind"] == "contain"]
    print("batched workload: captured batch, one replay", flush=True)
    for q in queries:
        if q["kind"] == "contain":
            q["out"].fill_(-7)
    cb = CapturedBatch(ctx, queries)
    try:
        graph_pairs = cb.run()
        graph_ids = [to_host(q["out"]).copy() for q in queries if q["kind"] == "contain"]
    finally:
        cb.close()
    print("batched workload: per-query launches and the oracle", flush=True)
    ci = 0
    n_pairs = 0
    for q, src, ep, gp in zip(queries, qs, eager_pairs, graph_pairs):
        if q["kind"] == "contain":
            alone = to_host(trace_points(ctx, q["bvh"], q["points"], axis=2))
            m = src["mesh"]
            ref = orc.points_contain(m.vertices, m.indices, m.tri_to_obj, src["points"], axis=2,
                                     threads=THREADS)
            assert np.array_equal(eager_ids[ci], alone) and np.array_equal(graph_ids[ci], alone), ci
            assert np.array_equal(alone, ref), f"query {ci}: {np.count_nonzero(alone != ref)} differ"
            assert len(ep) == 0 and len(gp) == 0
            ci += 1
        else:
            alone = overlap_join(ctx, q["bvh1"], q["edges1"], q["bvh2"], q["edges2"], mode=1)
            ref = orc.overlap_pairs(src["a"], src["a"].edges, src["b"], src["b"].edges)
            assert np.array_equal(ep, alone) and np.array_equal(gp, alone)
            assert np.array_equal(alone, ref)
            n_pairs += len(ref)
    assert n_pairs > 100
    for b in keep:
        b.free()


# ---- served form: 64 resident config-2 datasets, Poisson stream ---------------------------

def _dataset(i):
    mesh, _ = gen.many_spheres_workload(seed=1000 + i, n_objects=10_000, n_points=0)
    return mesh


def _obj_boxes(mesh):
    """Per-object boxes (the generators lay each object's triangles out contiguously, in
    object order)."""
    tv = mesh.vertices[mesh.indices.reshape(-1)].reshape(-1, 3, 3)
    first = np.searchsorted(mesh.tri_to_obj, np.arange(mesh.num_objects))
    assert np.all(np.diff(mesh.tri_to_obj) >= 0)
    return np.minimum.reduceat(tv.min(1), first), np.maximum.reduceat(tv.max(1), first)


def _near_subset(ds, ds_lo, ds_hi, qm):
    """The objects of dataset ds whose xy box overlaps the xy box of some object of the
    request mesh qm (widened by 1e-3). Every hit of the intersection join's four parts lies
    on such an object: an edge hit needs the two objects' boxes to meet, and a +z ray from a
    first vertex only meets objects whose xy box holds the vertex. Object ids are kept."""
    q_lo, q_hi = _obj_boxes(qm)
    keep = np.zeros(ds.num_objects, bool)
    for lo, hi in zip(q_lo, q_hi):
        keep |= ((ds_lo[:, 0] <= hi[0] + 1e-3) & (ds_hi[:, 0] >= lo[0] - 1e-3) &
                 (ds_lo[:, 1] <= hi[1] + 1e-3) & (ds_hi[:, 1] >= lo[1] - 1e-3))
    sel = keep[ds.tri_to_obj]
    sub = gen.Mesh(ds.vertices, np.ascontiguousarray(ds.indices[sel]), np.ascontiguousarray(ds.tri_to_obj[sel]),
                   ds.num_objects)
    sub.edges = gen.extract_edges(sub)
    return sub


def _serve_all(srv, reqs, ids, qids, pool_dev, outs, rng, burst):
    """Submits the requests in order, stepping the server after a Poisson-distributed
    number of submissions (burst = mean); returns ticket -> request index."""
    tickets = {}
    pending = 0
    next_step = rng.poisson(burst)
    for i, r in enumerate(reqs):
        if r[0] == _lib.REQ_CONTAIN:
            s, n = r[3], r[4]
            t = srv.submit(_lib.REQ_CONTAIN, ids[r[1]], points=pool_dev[s: s + n], out=outs[i])
        else:
            t = srv.submit(r[0], ids[r[1]], qids[r[2]])
        tickets[t] = i
        pending += 1
        if pending > next_step:
            srv.step()
            pending = 0
            next_step = rng.poisson(burst)
    while srv.stats()["queued"]:
        srv.step()
    return tickets


@pytest.mark.timeout(1200)
def test_serving_workload_64_datasets_poisson_equals_alone_and_oracle(ctx):
    rng = np.random.Generator(np.random.PCG64(55))
    n_ds = 64
    print(f"serving: building {n_ds} resident datasets", flush=True)
    bvhs, edges = [], []
    for i in range(n_ds):
        m = _dataset(i)
        bvhs.append(Bvh.from_mesh(ctx, m))
        edges.append(DevEdges(gen.sphere_field_edges(m, 2)).spatial(ctx))
    qmeshes, qbvh, qedges = [], [], []
    for i in range(12):
        k = int(np.exp(rng.uniform(0.0, np.log(64.0))))
        qm = gen.overlap_dataset(5000 + i, n_objects=max(k, 1), world=1000.0)
        qm.edges = gen.extract_edges(qm)
        qmeshes.append(qm)
        qbvh.append(Bvh.from_mesh(ctx, qm))
        qedges.append(DevEdges(qm.edges).spatial(ctx))
    pool = gen.uniform_points(rng, 4_000_000, 0.0, 1000.0)
    pool_dev = to_dev(pool)
    # 96 requests: 1/2 containment (1e3..1e6 points), 1/4 overlap, 1/4 intersection
    n_req = 96
    kinds = rng.choice([0, 0, 1, 2], size=n_req)
    targets = rng.integers(0, n_ds, size=n_req)
    qsel = rng.integers(0, len(qmeshes), size=n_req)
    npts = np.exp(rng.uniform(np.log(1e3), np.log(1e6), size=n_req)).astype(np.int64)
    starts = rng.integers(0, len(pool) - 1_000_000, size=n_req)
    reqs = [(int(kinds[i]), int(targets[i]), int(qsel[i]), int(starts[i]), int(npts[i])) for i in range(n_req)]

    results = {}
    for tag, slots, burst in (("mixed", 64, 6.0), ("alone", 1, 0.0)):
        print(f"serving: {tag} run of {n_req} requests", flush=True)
        # 40 GiB: room for a few axis grids (about 3 GB each, bounded at 6.3 GB while
        # building), so some containment requests run on a grid and others on the walk
        srv = Server(ctx, max_slots=slots, memory_budget=40 << 30)
        ids = [srv.add_dataset(b, e, grid_res=128) for b, e in zip(bvhs, edges)]
        qids = [srv.add_dataset(b, e, grid_res=128) for b, e in zip(qbvh, qedges)]
        outs = [torch.full((r[4],), -9, dtype=torch.int32, device="cuda") if r[0] == 0 else None for r in reqs]
        tickets = _serve_all(srv, reqs, ids, qids, pool_dev, outs, np.random.Generator(np.random.PCG64(56)), burst)
        res = {}
        for t, i in tickets.items():
            done, keys = srv.result(t)
            assert done == 1, (tag, i)
            res[i] = to_host(outs[i]).copy() if reqs[i][0] == 0 else np.asarray(keys, np.uint64)
        st = srv.stats()
        srv.close()
        results[tag] = res
        print(f"serving: {tag}: {st['steps']} steps, peak {st['peak_memory'] / 2**30:.1f} GiB", flush=True)
        assert st["peak_memory"] <= 40 << 30 or slots == 1
    for i in range(n_req):
        assert np.array_equal(results["mixed"][i], results["alone"][i]), i
    built = sum(1 for b in bvhs if _grid_state(ctx, b) == 1)
    print(f"serving: axis grids built for {built} of {n_ds} datasets (the rest on the walk)", flush=True)

    print("serving: oracle", flush=True)
    by_ds = {}
    for i, r in enumerate(reqs):
        if r[0] == 0:
            by_ds.setdefault(r[1], []).append(i)
    for d, idx in by_ds.items():
        m = _dataset(d)
        pts = np.concatenate([pool[reqs[i][3]: reqs[i][3] + reqs[i][4]] for i in idx])
        ref = orc.points_contain(m.vertices, m.indices, m.tri_to_obj, pts, axis=2, threads=THREADS)
        o = 0
        for i in idx:
            n = reqs[i][4]
            assert np.array_equal(results["mixed"][i], ref[o: o + n]), (i, d)
            o += n
    boxes = {}
    n_pairs = 0
    for i, r in enumerate(reqs):
        if r[0] == 0:
            continue
        d, qm = r[1], qmeshes[r[2]]
        if d not in boxes:
            m = _dataset(d)
            boxes[d] = (m,) + _obj_boxes(m)
        m, lo, hi = boxes[d]
        sub = _near_subset(m, lo, hi, qm)
        if r[0] == _lib.REQ_OVERLAP:
            ref = orc.overlap_pairs(qm, qm.edges, sub, sub.edges)
        else:
            ref = orc.intersection_pairs(qm, sub)
        assert np.array_equal(np.sort(results["mixed"][i]), ref), (i, len(results["mixed"][i]), len(ref))
        n_pairs += len(ref)
    assert n_pairs > 0
    for b in bvhs + qbvh:
        b.free()


def _grid_state(ctx, bvh, axis=2):
    import ctypes as C
    info = _lib.AxisGridInfo()
    _lib.check(ctx.lib.rs3d_axis_grid_query(bvh.handle, axis, C.byref(info)), "rs3d_axis_grid_query")
    return info.state


def test_server_budget_below_one_grid_builds_none(ctx):
    rng = np.random.Generator(np.random.PCG64(71))
    mesh, _ = gen.many_spheres_workload(seed=72, n_objects=2000, n_points=0, world=200.0)
    mesh.edges = gen.extract_edges(mesh)
    qm = gen.overlap_dataset(73, n_objects=30, world=200.0, levels=(2, 3))
    qm.edges = gen.extract_edges(qm)
    bd, bq = Bvh.from_mesh(ctx, mesh), Bvh.from_mesh(ctx, qm)
    ed, eq = DevEdges(mesh.edges), DevEdges(qm.edges)
    # 64 MiB: the first-triangle lists and 32^3 histograms fit, an axis grid (bounded at
    # about 96 B per record of 16 per triangle) or occupancy bits (bounded at 144 MiB) do not
    srv = Server(ctx, max_slots=8, memory_budget=64 << 20)
    D = srv.add_dataset(bd, ed, grid_res=32)
    Q = srv.add_dataset(bq, eq, grid_res=32)
    pts = [gen.uniform_points(rng, n, 0.0, 200.0) for n in (50_000, 3_000, 200_000)]
    dps = [to_dev(p) for p in pts]
    outs = [torch.full((len(p),), -9, dtype=torch.int32, device="cuda") for p in pts]
    ts = [srv.submit(_lib.REQ_CONTAIN, D, points=dp, out=o) for dp, o in zip(dps, outs)]
    t_ov = srv.submit(_lib.REQ_OVERLAP, D, Q)
    t_in = srv.submit(_lib.REQ_INTERSECT, D, Q)
    while srv.stats()["queued"]:
        srv.step()
    assert _grid_state(ctx, bd) == 0  # never built: every containment request walked the BVH
    for t, p, o in zip(ts, pts, outs):
        done, _ = srv.result(t)
        assert done == 1
        ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, p, axis=2, threads=THREADS)
        assert np.array_equal(to_host(o), ref)
    done, keys = srv.result(t_ov)
    assert done == 1 and np.array_equal(np.sort(np.asarray(keys, np.uint64)),
                                        orc.overlap_pairs(qm, qm.edges, mesh, mesh.edges))
    done, keys = srv.result(t_in)
    ref = orc.intersection_pairs(qm, mesh)
    assert done == 1 and np.array_equal(np.sort(np.asarray(keys, np.uint64)), ref) and len(ref) > 0
    srv.close()
    bd.free()
    bq.free()
