"""The query executables (C++ over the C ABI) on generated .pre inputs: CSV contents equal
the oracle's pair sets and containment ids; timing JSON has the reference's key scheme."""
import csv
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from rayspace3d_amd import gen, prep, preio

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rayspace3d_amd", "bin")


def run(exe, args, cwd):
    print(f"running {exe} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(BIN, exe)] + args, cwd=cwd, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def read_pairs(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], [tuple(int(x) for x in r[:2]) + tuple(r[2:]) for r in rows[1:]]


def keys_to_set(keys):
    return {(int(k >> np.uint64(32)), int(k & np.uint64(0xffffffff))) for k in keys}


@pytest.fixture(scope="module")
def datasets(tmp_path_factory, ctx):
    d = tmp_path_factory.mktemp("cli")
    rng = np.random.Generator(np.random.PCG64(77))
    n = 150
    ca = rng.uniform(10, 140, size=(n, 3))
    a = gen.spheres(rng.choice([2, 3], size=n), ca, np.full(n, 5.0), rotations=gen.random_rotations(rng, n))
    cb = np.concatenate([ca[:60] + rng.uniform(-1, 1, size=(60, 3)), rng.uniform(10, 140, size=(n - 60, 3))])
    b = gen.spheres(2, cb, np.concatenate([np.full(60, 2.0), np.full(n - 60, 3.5)]),
                    rotations=gen.random_rotations(rng, n))
    prep.write_dataset(str(d / "a.pre"), a, ctx, grid_res=16, world_size=150.0)
    prep.write_dataset(str(d / "b.pre"), b, ctx, grid_res=16, world_size=150.0)
    pts = gen.uniform_points(rng, 30000, 0.0, 150.0)
    prep.write_points_wkt(str(d / "pts.wkt"), pts)
    pts_back = np.array([[np.float32(x) for x in l.split("(")[1].split(")")[0].split()] for l in open(d / "pts.wkt")],
                        np.float32)
    return d, a, b, pts_back


def test_raytracer_and_filter_refine(datasets):
    d, a, b, pts = datasets
    ref = orc.points_contain(a.vertices, a.indices, a.tri_to_obj, pts, axis=2)
    exp = {(i, int(o)) for i, o in enumerate(ref) if o >= 0}
    run("raytracer", ["--geometry", "a.pre", "--points", "pts.wkt", "--warmup-runs", "1"], d)
    hdr, rows = read_pairs(d / "ray_results.csv")
    assert hdr == ["pointId", "polygonId"] and set(rows) == exp
    t = json.load(open(d / "performance_timing.json"))
    for k in ("load geometry_1", "build index_1", "query_1", "output_1"):
        assert k in t["phases"]
    run("raytracer_filter_refine", ["--geometry", "a.pre", "--points", "pts.wkt"], d)
    hdr, rows = read_pairs(d / "ray_results.csv")
    assert set(rows) == exp


def test_mesh_overlap_two_pass_order(datasets):
    d, a, b, _ = datasets
    ref = orc.overlap_pairs(a, a.edges, b, b.edges)
    run("raytracer_mesh_overlap", ["--mesh1", "a.pre", "--mesh2", "b.pre", "--warmup-runs", "1"], d)
    hdr, rows = read_pairs(d / "mesh_overlap_results.csv")
    assert hdr == ["object_id_mesh1", "object_id_mesh2"]
    assert set(rows) == keys_to_set(ref) and len(rows) == len(ref)
    assert rows == sorted(rows, key=lambda r: (r[1], r[0]))  # reference (b, a) order
    t = json.load(open(d / "mesh_overlap_timing.json"))
    assert "raytrace_mesh1tomesh2_pass1_1" in t["phases"]


def test_overlap_estimated_and_direct(datasets):
    d, a, b, _ = datasets
    ref = orc.overlap_pairs(a, a.edges, b, b.edges)
    out = run("raytracer_overlap_estimated", ["--mesh1", "a.pre", "--mesh2", "b.pre"], d)
    assert f"Unique object pairs: {len(ref)}" in out
    run("raytracer_overlap_direct_estimation", ["--mesh1", "a.pre", "--mesh2", "b.pre",
                                                "--pairs-output", "pairs.csv", "--hash-table-size", "1024"], d)
    _, rows = read_pairs(d / "pairs.csv")
    assert set(rows) == keys_to_set(ref)


def test_intersection_cli(datasets):
    d, a, b, _ = datasets
    ref = orc.intersection_pairs(a, b)
    for exe in ("raytracer_intersection_estimated", "raytracer_mesh_intersection"):
        run(exe, ["--mesh1", "a.pre", "--mesh2", "b.pre", "--enable-profiling-stats"], d)
        hdr, rows = read_pairs(d / "intersection_pairs.csv")
        assert hdr == ["a_object_id", "b_object_id"] and set(rows) == keys_to_set(ref)
    # the tracking CSVs (PairHitTracking.h, ContainmentTracking.h) against the oracle's tracking
    trk = {}
    for src, tgt, swap, tag in ((a, b, False, "mesh1_to_mesh2"), (b, a, True, "mesh2_to_mesh1")):
        _, _, tr, st = orc.object_parity_ex(src, tgt, 256, swap, 0, track=True, max_pair=256, max_out=256)
        trk[tag] = (tr, st)
    with open(d / "intersection_pair_hits.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["direction", "source_object_id", "target_object_id", "target_ray_hits"]
    exp = []
    for tag in ("mesh1_to_mesh2", "mesh2_to_mesh1"):
        tr = trk[tag][0]
        for s_, (ids, hits) in enumerate(zip(tr["pair_target_ids"], tr["pair_target_hits"])):
            exp += [[tag, str(s_), str(int(t)), str(int(h))] for t, h in zip(ids, hits) if t >= 0]
    assert rows[1:] == exp and len(exp) > 10
    with open(d / "intersection_containment_tracking.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["direction", "source_object_id", "iterations", "candidate_count", "candidate_overflow_events"]
    exp = []
    for tag in ("mesh1_to_mesh2", "mesh2_to_mesh1"):
        tr = trk[tag][0]
        exp += [[tag, str(s_), str(int(i)), str(int(c)), str(int(o))]
                for s_, (i, c, o) in enumerate(zip(tr["iterations"], tr["candidate_count"], tr["candidate_overflow"]))]
    assert rows[1:] == exp
    # profiling counters in the timing JSON (raytracer_intersection_estimated_main.cpp:722-747)
    cnt = json.load(open(d / "mesh_intersection_timing.json"))["counters"]
    st = [trk[k][1] for k in trk]
    for name, key in (("profile_containment_rays_total", "rays"), ("profile_containment_hits_total", "hits"),
                      ("profile_containment_candidate_toggles", "toggles"),
                      ("profile_containment_targets_total", "targets")):
        assert cnt[name] == st[0][key] + st[1][key], name
    assert cnt["profile_overlap_trace_calls"] > 0 and cnt["profile_actual_intersection_pairs"] == len(ref)


def test_containment_cli(datasets):
    d, a, b, _ = datasets
    ov, ct = orc.containment_pairs(a, b)
    run("raytracer_containment", ["--mesh1", "a.pre", "--mesh2", "b.pre", "--include-overlap-pairs"], d)
    hdr, rows = read_pairs(d / "mesh_containment_results.csv")
    assert hdr == ["a_object_id", "b_object_id", "pair_type"]
    assert {r[:2] for r in rows if r[2] == "containment"} == keys_to_set(ct)
    assert {r[:2] for r in rows if r[2] == "overlap"} == keys_to_set(ov)
    assert len(ct) >= 40


def _obj_text(mesh):
    """OBJ of a mesh: one 'o' group per object holding the vertices it uses."""
    out = []
    nxt = 1
    for k in range(int(mesh.tri_to_obj.max()) + 1):
        tris = mesh.indices[mesh.tri_to_obj == k]
        used = np.unique(tris)
        out.append(f"o object_{k}")
        out += [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in mesh.vertices[used].astype(np.float64)]
        pos = {int(v): nxt + i for i, v in enumerate(used)}
        out += ["f " + " ".join(str(pos[int(v)]) for v in t) for t in tris]
        nxt += len(used)
    return "\n".join(out) + "\n"


def _assert_edges(pre_edges, V, I, O):
    S, E, off, cnt, src = orc.extract_edges(V, I, O)
    assert np.array_equal(pre_edges.starts.view(np.uint32), S.view(np.uint32))
    assert np.array_equal(pre_edges.ends.view(np.uint32), E.view(np.uint32))
    assert np.array_equal(pre_edges.src_offsets, off) and np.array_equal(pre_edges.src_counts, cnt)
    assert np.array_equal(pre_edges.src_objects, src)


def test_preprocess_dataset(datasets):
    import hostio
    from rayspace3d_amd import preio
    d, a, b, pts = datasets
    text = _obj_text(a)
    (d / "a.obj").write_text(text)
    V, I, O = hostio.load_obj(text)
    assert len(I) == len(a.indices)
    run("preprocess_dataset", ["--dataset", "a.obj", "--output-geometry", "a_obj.pre", "--generate-grid",
                               "--grid-resolution", "16", "--world-size", "150", "--output-timing", "pp.json"], d)
    pre = preio.read_pre(str(d / "a_obj.pre"))
    assert pre.version == 2
    assert np.array_equal(pre.mesh.vertices.view(np.uint32), V.view(np.uint32))
    assert np.array_equal(pre.mesh.indices, I) and np.array_equal(pre.mesh.tri_to_obj, O)
    _assert_edges(pre.mesh.edges, V, I, O)
    cells, mn, mx = orc.grid_stats(V, I, O, 16, 150.0)
    g = pre.grid
    assert g.resolution == (16, 16, 16) and np.allclose(g.min_bound, mn) and np.allclose(g.max_bound, mx)
    assert np.array_equal(g.cells["center"], cells["center"]) and np.array_equal(g.cells["touch"], cells["touch"])
    assert np.allclose(g.cells["avg"], cells["avg"], rtol=1e-5)
    assert np.allclose(g.cells["ratio"], cells["ratio"], rtol=1e-4)
    t = json.load(open(d / "pp.json"))
    for k in ("loading dataset_1", "generating grid statistics_1", "extracting edges_1", "writing geometry data_1"):
        assert k in t["phases"], k
    # v1 -> v2: the reference's committed v1 fixture gains the edge payload, grid unchanged
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "check_n_geometry.pre")
    run("preprocess_dataset", ["--mode", "pre", "--dataset", golden, "--output-geometry", "g2.pre",
                               "--output-timing", "pp2.json"], d)
    g1, g2 = preio.read_pre(golden), preio.read_pre(str(d / "g2.pre"))
    assert g1.version == 1 and g2.version == 2
    m = g1.mesh
    assert np.array_equal(g2.mesh.vertices, m.vertices) and np.array_equal(g2.mesh.indices, m.indices)
    _assert_edges(g2.mesh.edges, m.vertices, m.indices, m.tri_to_obj)
    assert g2.grid.resolution == g1.grid.resolution and g2.grid.cells.tobytes() == g1.grid.cells.tobytes()
    # the converted .pre runs through the query executables; binary points give the WKT answer
    run("preprocess_dataset", ["--mode", "points", "--dataset", "pts.wkt", "--output-geometry", "pts.rs3p",
                               "--output-timing", "pp3.json"], d)
    ref = orc.points_contain(V, I, O, pts, axis=2)
    exp = {(i, int(o)) for i, o in enumerate(ref) if o >= 0}
    run("raytracer", ["--geometry", "a_obj.pre", "--points", "pts.rs3p", "--warmup-runs", "0"], d)
    hdr, rows = read_pairs(d / "ray_results.csv")
    assert set(rows) == exp


def test_mesh_overlap_exact(datasets):
    d, a, b, pts = datasets
    ref = orc.tri_join(a, b, threads=8)
    exp = sorted(keys_to_set(ref), key=lambda p: (p[1], p[0]))
    run("raytracer_mesh_overlap", ["--mesh1", "a.pre", "--mesh2", "b.pre", "--exact", "--warmup-runs", "1"], d)
    hdr, rows = read_pairs(d / "mesh_overlap_results.csv")
    assert hdr == ["object_id_mesh1", "object_id_mesh2"] and rows == exp and len(rows) > 10
    # the exact join needs no edge payload: a v1 file without edges is accepted
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "check_n_geometry.pre")
    run("raytracer_mesh_overlap", ["--mesh1", golden, "--mesh2", golden, "--exact", "--warmup-runs", "0"], d)
    hdr, rows = read_pairs(d / "mesh_overlap_results.csv")
    assert len(rows) >= 540 and all((i, i) in set(rows) for i in range(540))


def _read_launchers(path):
    sets, stats, est, flags = {}, {}, {}, {}
    for line in open(path):
        w = line.split()
        if w[0] == "stat":
            stats[w[1]] = int(w[2])
        elif w[0] == "estimate":
            est[w[1]] = float(w[2])
        elif len(w) == 2:
            flags[w[0]] = int(w[1])
        else:
            n = int(w[1])
            sets[w[0]] = np.array([int(x) for x in w[2:2 + n]], np.uint64)
    return sets, stats, est, flags


def test_launcher_classes_and_estimators(datasets):
    """MeshIntersectionLauncher (every method, any-hit and closest-hit containment, hash and
    two-pass forms, profiling counters), MeshContainmentLauncher (edge check and point in mesh,
    both point-in-mesh forms) and estimateOverlapSelectivity / estimateIntersectionSelectivity
    called from C++ (check_launchers), against the oracle."""
    d, a, b, _ = datasets
    run("check_launchers", ["--mesh1", "a.pre", "--mesh2", "b.pre", "--out", "launchers.txt"], d)
    sets, stats, est, flags = _read_launchers(d / "launchers.txt")
    # intersection launcher
    assert np.array_equal(sets["intersection_anyhit"], orc.intersection_pairs(a, b))
    k1 = orc.edge_pairs(a.edges, b.vertices, b.indices, b.tri_to_obj, False)
    k2 = orc.edge_pairs(b.edges, a.vertices, a.indices, a.tri_to_obj, True)
    c1 = orc.object_parity_ex(a, b, 256, False, 1, 1024, max_out=256)[0]
    c2 = orc.object_parity_ex(b, a, 256, True, 1, 1024, max_out=256)[0]
    assert np.array_equal(sets["intersection_closest"], orc.sort_unique(np.concatenate([k1, k2, c1, c2])))
    ov12 = orc.sort_unique(k1)
    assert np.array_equal(sets["overlap12_hash"], ov12) and np.array_equal(sets["overlap12_twopass"], ov12)
    assert flags["merged_in_ba_order"] == 1 and len(ov12) > 5
    ct12, _, _, st1 = orc.object_parity_ex(a, b, 256, False, 0, max_out=256)
    assert np.array_equal(sets["contain12_hash"], orc.sort_unique(ct12))
    assert np.array_equal(sets["contain12_twopass"], orc.sort_unique(ct12)) and len(ct12) > 0
    # profiling counters of the any-hit run (both directions)
    st2 = orc.object_parity_ex(b, a, 256, True, 0, max_out=256)[3]
    for name, key in (("containment_rays_total", "rays"), ("containment_iterations_total", "iterations"),
                      ("containment_hits_total", "hits"), ("containment_candidate_additions", "additions"),
                      ("containment_candidate_toggles", "toggles"), ("containment_targets_total", "targets")):
        assert stats[name] == st1[key] + st2[key], name
    calls = iters = hits = mx = 0
    for e, m in ((a.edges, b), (b.edges, a)):
        _, cnt = orc.edge_pairs(e, m.vertices, m.indices, m.tri_to_obj, with_counts=True)
        dv = (e.ends - e.starts).astype(np.float32)
        ln = np.sqrt((dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1]) + dv[:, 2] * dv[:, 2]).astype(np.float32)
        live = (ln >= np.float32(1e-6)) & (e.src_counts > 0)
        sc = e.src_counts.astype(np.int64)
        it = np.minimum(np.where(live, cnt // np.maximum(sc, 1), 0) + 1, 100)
        calls += int(sc[live].sum())
        iters += int((it * sc)[live].sum())
        hits += int(cnt[live].sum())
        mx = max(mx, int(it[live].max()))
    assert (stats["overlap_trace_calls"], stats["overlap_iterations_total"], stats["overlap_hits_total"],
            stats["overlap_max_iterations_per_trace"]) == (calls, iters, hits, mx)
    # containment launcher: any-hit form equals the containment join; the closest-hit form
    # steps t + 1e-4 from B's first vertices
    ov, ct = orc.containment_pairs(a, b)
    assert np.array_equal(sets["containment_overlap"], ov)
    assert np.array_equal(sets["containment_anyhit"], ct) and len(ct) >= 40
    first = np.stack([b.vertices[b.indices[np.nonzero(b.tri_to_obj == o)[0][0], 0]] for o in range(b.num_objects)])
    cand = orc.sort_unique(orc.object_parity_ex(None, a, 512, True, 2, 2000, origins=first)[0])
    assert np.array_equal(sets["containment_closest"], cand[~np.isin(cand, ov)])
    # estimators (reference signatures): raw estimate, and with the host alpha correction
    ga, gb = preio.read_pre(str(d / "a.pre")).grid, preio.read_pre(str(d / "b.pre")).grid
    cv = float(np.prod((ga.max_bound.astype(np.float32) - ga.min_bound.astype(np.float32)) / np.float32(16)))
    raw = orc.estimate_raw(ga.cells, gb.cells, cv)
    assert abs(est["overlap_raw"] - raw) <= 1e-5 * raw and abs(est["intersection_raw"] - raw) <= 1e-5 * raw
    assert est["empty"] == 0.0
    assert est["pairs"] == orc.estimate_pairs(ga.cells, gb.cells, ga.min_bound, ga.max_bound, ga.resolution)
