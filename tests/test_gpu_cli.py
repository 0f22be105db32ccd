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
