"""The query executables (C++ over the C ABI) on generated .pre inputs: CSV contents equal
the oracle's pair sets and containment ids; timing JSON has the reference's key scheme."""
import csv
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from rayspace3d_amd import gen, prep

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
        run(exe, ["--mesh1", "a.pre", "--mesh2", "b.pre"], d)
        hdr, rows = read_pairs(d / "intersection_pairs.csv")
        assert hdr == ["a_object_id", "b_object_id"] and set(rows) == keys_to_set(ref)


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
