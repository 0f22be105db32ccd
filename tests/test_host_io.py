"""Dataset input of the host layer, on the CPU: the restated reference loaders (oracle/hostio.py)
pinned on hand-computed answers, then the C++ loaders (preprocess_dataset --mode points /
--mode mesh --skip-edges, which use no GPU) checked against them."""
import os
import struct
import subprocess

import numpy as np
import pytest

import hostio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "rayspace3d_amd", "bin", "preprocess_dataset")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="preprocess_dataset not built")


def run_pre(args, cwd):
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    return r


def read_points_bin(path):
    b = open(path, "rb").read()
    magic, ver, n, stride = struct.unpack("<IIQI", b[:20])
    assert (magic, ver, stride) == (0x50335352, 1, 12) and len(b) == 32 + 12 * n
    return np.frombuffer(b[32:], np.float32).reshape(-1, 3)


def read_pre(path):
    from rayspace3d_amd import preio
    return preio.read_pre(path)


# ---- oracle pins ------------------------------------------------------------------------

def test_point_parser_known_answers():
    data = (b"POINT Z (1 2 3)\n"
            b"  POINT Z (-1.5 0.25 1e2)\r\n"          # leading blanks, CRLF
            b"# comment 4 5 6\n\n\t\n"
            b"4 5 6\n"
            b"-7 8\n"                                  # missing z -> strtof gives 0
            b"+1 2 3\n"                                # '+' is not a start character
            b"POINT Z 1 2 3\n"                         # no '(' -> skipped
            b"POINT EMPTY ()\n"                        # '(' present -> (0, 0, 0)
            b"x 1 2\n"
            b"9 10 11")                                # no trailing newline
    got = hostio.parse_points(data)
    exp = np.array([[1, 2, 3], [-1.5, 0.25, 100], [4, 5, 6], [-7, 8, 0], [0, 0, 0], [9, 10, 11]], np.float32)
    np.testing.assert_array_equal(got, exp)


def test_point_parser_long_line_pieces():
    # fgets with a 1024-byte buffer hands a long line over in 1023-byte pieces; a piece that
    # starts with a digit is parsed as a line of its own.
    pad = b" " * 1020
    line = pad + b"1 2 3 " + b"4 5 6\n"               # piece 1 = pad + "1 2", piece 2 = " 3 4 5 6\n"
    got = hostio.parse_points(line)
    np.testing.assert_array_equal(got, np.array([[1, 2, 0], [3, 4, 5]], np.float32))
    assert len(hostio.parse_points(b" " * 2000 + b"1 2 3\n")) == 1   # pieces of blanks are ignored


def test_tinyobj_number_parser():
    # Short decimals: the digit-by-digit sum rounds to float exactly like strtof.
    rng = np.random.default_rng(5)
    for _ in range(2000):
        s = f"{rng.uniform(-500, 500):.{rng.integers(0, 7)}f}"
        assert np.float32(hostio.tobj_parse_double(s)) == np.float32(float(s)), s
    assert hostio.tobj_parse_double("1e3") == 1000.0
    assert hostio.tobj_parse_double("-.5") == -0.5
    assert hostio.tobj_parse_double("2.5E-1") == 0.25
    assert hostio.tobj_parse_double("7.") == 7.0
    assert hostio.tobj_parse_double("1.5abc") == 1.5
    assert hostio.tobj_parse_double(".e1") == 0.0             # no digits at all, still a number
    for bad in ("", "abc", "-", "1e", "1e+", "e5"):
        assert hostio.tobj_parse_double(bad) is None, bad
    # A 16-digit decimal just above a float rounding midpoint: the double sum of the
    # fraction digits lands below it, so the loader's float is one ulp below strtof's.
    s = "4.614706277847291"
    strtof = np.float32(float(s))
    assert np.float32(hostio.tobj_parse_double(s)) == np.nextafter(strtof, np.float32(0))


CUBE = """# unit cube, 6 quads, one object
o cube
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
f 1 2 3 4
f 5 8 7 6
f 1 5 6 2
f 2 6 7 3
f 3 7 8 4
f 5 1 4 8
"""


def test_obj_known_answers():
    V, I, O = hostio.load_obj(CUBE)
    assert V.shape == (8, 3) and I.shape == (12, 3) and (O == 0).all()
    # Square faces: both diagonals are equal (d02 < d13 false) -> [0,1,3], [1,2,3].
    assert I[:2].tolist() == [[0, 1, 3], [1, 2, 3]]
    # A long thin quad splits along its short diagonal.
    q = "v 0 0 0\nv 10 0 0\nv 10.5 1 0\nv 0 1 0\nf 1 2 3 4\n"
    V, I, O = hostio.load_obj(q)
    # d02 = 10.5^2 + 1 = 111.25, d13 = 10^2 + 1 = 101 -> split along 1-3
    assert I.tolist() == [[0, 1, 3], [1, 2, 3]]
    q2 = "v 0 0 0\nv 10.5 0 0\nv 10 1 0\nv 0 1 0\nf 1 2 3 4\n"
    assert hostio.load_obj(q2)[1].tolist() == [[0, 1, 2], [0, 2, 3]]
    # Groups, relative indices, slashes, a lines-only group, an out-of-range vertex.
    text = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
            "g a\nf 1/1/1 2/1/1 3/1/1\n"
            "g lines\nl 1 2\n"
            "o b\nv 5 5 5\nv 6 5 5\nv 5 6 5\nf -3 -2 -1\nf 4//1 5//1 9//1\n"
            "g\nf 1 2 3\n")                     # bare 'g' is not a group line
    V, I, O = hostio.load_obj(text)
    # Object b = file vertices {0,1,2,3,4,5} in ascending order (8 is missing, its triangle
    # dropped); the face after the bare 'g' still belongs to b.
    assert O.tolist() == [0, 1, 1]              # lines-only group skipped
    assert I.tolist() == [[0, 1, 2], [6, 7, 8], [3, 4, 5]]
    np.testing.assert_array_equal(V[6:9], [[5, 5, 5], [6, 5, 5], [5, 6, 5]])
    np.testing.assert_array_equal(V[3:6], V[:3])  # vertices shared across objects are copied
    # An index of 0, or a relative index before the first vertex, fails the whole load.
    assert hostio.load_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n") is None
    assert hostio.load_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nf -4 1 2\n") is None
    # An object whose only triangle references missing vertices is still counted.
    V, I, O = hostio.load_obj("o x\nv 0 0 0\nf 1 7 8\no y\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    assert O.tolist() == [1] and I.tolist() == [[1, 2, 3]]


# ---- C++ loaders vs the oracle ------------------------------------------------------------

def _random_point_text(rng, n):
    lines = []
    for i in range(n):
        x, y, z = rng.uniform(-1000, 1000, 3)
        k = rng.integers(0, 9)
        if k == 0:
            lines.append(f"POINT Z ({x:.9g} {y:.9g} {z:.9g})")
        elif k == 1:
            lines.append(f"  POINT Z ({x:.3f} {y:.3f} {z:.3f})\r")
        elif k == 2:
            lines.append(f"{x:.7e} {y:.7e} {z:.7e}")
        elif k == 3:
            lines.append("# " + "c" * int(rng.integers(0, 50)))
        elif k == 4:
            lines.append("")
        elif k == 5:
            lines.append(" " * int(rng.integers(900, 1100)) + f"{x:.4f} {y:.4f} {z:.4f}")
        elif k == 6:
            lines.append(f"\t{x:.17g}\t{y:.17g}")
        elif k == 7:
            lines.append(f"POINT ({x} {y} {z}) trailing")
        else:
            lines.append(f"{abs(x):.2f} {y:.2f} {z:.2f}")
    return ("\n".join(lines)).encode()


@needs_exe
def test_points_text_and_binary(tmp_path):
    rng = np.random.default_rng(11)
    data = _random_point_text(rng, 40000)            # > 4 MiB: parsed by several workers
    (tmp_path / "p.wkt").write_bytes(data)
    exp = hostio.parse_points(data)
    r = run_pre(["--mode", "points", "--dataset", "p.wkt", "--output-geometry", "p.bin"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_points_bin(tmp_path / "p.bin")
    assert got.shape == exp.shape
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
    # Binary in, binary out: recognised by its magic, copied unchanged.
    r = run_pre(["--mode", "points", "--dataset", "p.bin", "--output-geometry", "q.bin"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "q.bin").read_bytes() == (tmp_path / "p.bin").read_bytes()
    # A truncated binary file is an error, not a short point set.
    (tmp_path / "t.bin").write_bytes((tmp_path / "p.bin").read_bytes()[:-4])
    r = run_pre(["--mode", "points", "--dataset", "t.bin", "--output-geometry", "u.bin"], tmp_path)
    assert r.returncode != 0


def _sphere_obj(rng, n_obj, quads=True):
    """OBJ text of UV spheres (quads + triangle caps) with random formatting."""
    out, nv = [], 0
    for o in range(n_obj):
        c = rng.uniform(0, 100, 3)
        r = rng.uniform(1, 5)
        nu, nl = int(rng.integers(4, 9)), int(rng.integers(3, 6))
        out.append(f"{'o' if o % 2 else 'g'} obj{o}")
        verts = [c + [0, 0, r]]
        for i in range(1, nl):
            th = np.pi * i / nl
            for j in range(nu):
                ph = 2 * np.pi * j / nu
                verts.append(c + r * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]))
        verts.append(c + [0, 0, -r])
        for p in verts:
            fmt = ["{:.6f}", "{:.9g}", "{:.15g}", "{:.4e}"][int(rng.integers(0, 4))]
            out.append("v " + " ".join(fmt.format(x) for x in p))
        base = nv + 1
        nv += len(verts)
        rel = bool(rng.integers(0, 2))

        def ref(k):
            return str(k - (nv + 1)) if rel else str(k)
        ring = lambda i, j: base + 1 + (i - 1) * nu + (j % nu)
        for j in range(nu):
            out.append(f"f {ref(base)} {ref(ring(1, j))}/1 {ref(ring(1, j + 1))}//1")
        for i in range(1, nl - 1):
            for j in range(nu):
                q = [ring(i, j), ring(i + 1, j), ring(i + 1, j + 1), ring(i, j + 1)]
                if quads:
                    out.append("f " + " ".join(ref(k) for k in q))
                else:
                    out.append(f"f {ref(q[0])} {ref(q[1])} {ref(q[2])}")
                    out.append(f"f {ref(q[0])} {ref(q[2])} {ref(q[3])}")
        last = base + len(verts) - 1
        for j in range(nu):
            out.append(f"f {ref(last)} {ref(ring(nl - 1, j + 1))} {ref(ring(nl - 1, j))}")
        out.append("vt 0.5 0.5\nvn 0 0 1")
    return "\r\n".join(out) + "\n"


@needs_exe
@pytest.mark.parametrize("quads", [True, False])
def test_obj_loader_matches_oracle(tmp_path, quads):
    rng = np.random.default_rng(21 + quads)
    text = _sphere_obj(rng, 60, quads)
    (tmp_path / "m.obj").write_text(text)
    exp = hostio.load_obj(text)
    r = run_pre(["--dataset", "m.obj", "--output-geometry", "m.pre", "--skip-edges"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    pre = read_pre(tmp_path / "m.pre")
    np.testing.assert_array_equal(pre.mesh.vertices.view(np.uint32), exp[0].view(np.uint32))
    np.testing.assert_array_equal(pre.mesh.indices, exp[1])
    np.testing.assert_array_equal(pre.mesh.tri_to_obj, exp[2])
    assert pre.version == 2 and pre.mesh.edges is None and pre.total_triangles == len(exp[1])


@needs_exe
def test_obj_loader_errors(tmp_path):
    (tmp_path / "z.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n")
    assert run_pre(["--dataset", "z.obj", "--output-geometry", "z.pre", "--skip-edges"], tmp_path).returncode == 1
    (tmp_path / "e.obj").write_text("# nothing\n")
    assert run_pre(["--dataset", "e.obj", "--output-geometry", "e.pre", "--skip-edges"], tmp_path).returncode == 1
    (tmp_path / "p.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nv 2 2 0\nf 1 2 3 4 5\n")
    assert run_pre(["--dataset", "p.obj", "--output-geometry", "p.pre", "--skip-edges"], tmp_path).returncode == 1
    assert run_pre(["--mode", "wkt", "--dataset", "x.wkt"], tmp_path).returncode == 1
    assert run_pre(["--dataset", "x.dt"], tmp_path).returncode == 1
