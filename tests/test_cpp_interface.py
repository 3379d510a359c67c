"""The C++ interface (include/ccdamd/, reference class and function signatures over the C ABI):
compiles and links without a GPU; on the GPU its results equal the Python front end's."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from collisiondetection_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_build", "cpp_interface_driver")


def _build_driver():
    subprocess.run(["bash", os.path.join(ROOT, "tests", "native", "build_cpp_driver.sh")], check=True)
    return DRIVER


def test_cpp_interface_compiles_and_links():
    assert os.path.exists(_build_driver())


@pytest.mark.gpu
def test_cpp_interface_matches_oracle(tmp_path):
    exe = DRIVER if os.path.exists(DRIVER) else _build_driver()
    s = scenes.converging_spheres()
    f, a, b = tmp_path / "f.i32", tmp_path / "a.f64", tmp_path / "b.f64"
    s["faces"].astype("<i4").tofile(f)
    s["qstart"].astype("<f8").tofile(a)
    s["qend"].astype("<f8").tofile(b)
    out = tmp_path / "out.txt"
    print("case: C++ driver on converging spheres", flush=True)
    subprocess.run([exe, str(f), str(a), str(b), repr(s["eta"]), str(out)], check=True, timeout=120)
    lines = out.read_text().splitlines()
    cand = [int(x) for x in lines[0].split()[1:]]
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
    assert cand == [len(vf), len(ee)]
    rv = oracle.narrowphase(h, vf, s["eta"], False)
    re = oracle.narrowphase(h, ee, s["eta"], True)
    got_vf = {tuple(int(x) for x in l.split()[1:]) for l in lines if l.startswith("vf ")}
    got_ee = {tuple(int(x) for x in l.split()[1:]) for l in lines if l.startswith("ee ")}
    for got, keys, rec in ((got_vf, vf, rv), (got_ee, ee, re)):
        o = {tuple(k) for k, hit in zip(keys.tolist(), rec["hit"]) if hit}
        border = {tuple(k) for k, m in zip(keys.tolist(), rec["margin"]) if m < 1e-12}
        assert got ^ o <= border
    # SeparatingPlaneNarrowPhase through the C++ class: the oracle's statuses (no undecided stencil here)
    for pre, keys, is_ee in (("spvf ", vf, False), ("spee ", ee, True)):
        got = {tuple(int(x) for x in l.split()[1:]) for l in lines if l.startswith(pre)}
        st, margin, _ = oracle.separating_plane(h, keys, s["eta"], is_ee)
        assert not (st == 2).any()
        o = {tuple(k) for k, x in zip(keys.tolist(), st) if x == 1}
        border = {tuple(k) for k, m in zip(keys.tolist(), margin) if m < 1e-12}
        assert got ^ o <= border
    prims = {l.split()[0]: (int(l.split()[1]), float(l.split()[2])) for l in lines if "_prim" in l}
    assert prims["vv_prim"] == (1, 0.1666665000090396)
    assert prims["ve_prim"] == (1, 0.24999975000276525)
    assert prims["vf_prim"][0] == 1 and abs(prims["vf_prim"][1] - 0.6666651759903873) < 1e-10
    assert prims["ee_prim"][0] == 1 and abs(prims["ee_prim"][1] - 0.5) < 1e-12


HISTORY_CHECK = os.path.join(ROOT, "tests", "_build", "history_check")


def _write_history_case(path, h, queries, stitches):
    with open(path, "wb") as f:
        np.array([h.nverts, len(h.ht)], np.int64).tofile(f)
        h.off.astype(np.int64).tofile(f)
        h.ht.astype(np.float64).tofile(f)
        h.hp.astype(np.float64).tofile(f)
        np.array([len(queries)], np.int64).tofile(f)
        for v, t in queries:
            np.array([v], np.int32).tofile(f)
            np.array([t], np.float64).tofile(f)
        np.array([len(stitches)], np.int64).tofile(f)
        for verts in stitches:
            np.array([len(verts)] + list(verts) + [0] * (4 - len(verts)), np.int32).tofile(f)


def test_history_methods_equal_oracle(tmp_path):
    """History::getPosAtTime (History.cpp:42-76) and History::stitchCommonHistory (History.cpp:98-140)
    of the C++ interface, the host methods a third-party narrowphase plugin calls: stitched times and
    positions equal the oracle's stitch bit for bit (which the oracle tests pin to the exact
    piecewise-linear paths); positions at time equal the reference's formula (1 - a) p + a q on the
    segment starting at the last entry at or before that time, and the last entry at or after t = 1."""
    from tests.test_oracle import _path_at, _random_histories

    exe = HISTORY_CHECK if os.path.exists(HISTORY_CHECK) else (_build_driver(), HISTORY_CHECK)[1]
    rng = np.random.default_rng(8)
    h = _random_histories(30, rng)
    queries = [(int(rng.integers(0, h.nverts)), float(t)) for t in rng.uniform(0, 1, size=300)]
    queries += [(v, float(h.ht[e])) for v in range(h.nverts) for e in range(h.off[v], h.off[v + 1])]
    stitches = [tuple(int(x) for x in rng.choice(h.nverts, size=int(rng.integers(1, 5)), replace=False))
                for _ in range(100)]
    case, out = tmp_path / "h.bin", tmp_path / "h.txt"
    _write_history_case(case, h, queries, stitches)
    subprocess.run([exe, str(case), str(out)], check=True, timeout=60)
    lines = out.read_text().splitlines()
    for (v, t), line in zip(queries, lines[:len(queries)]):
        f = line.split()
        idx, pos = int(f[1]), np.array([float.fromhex(x) for x in f[2:5]])
        ts = h.ht[h.off[v]:h.off[v + 1]]
        P = h.hp.reshape(-1, 3)[h.off[v]:h.off[v + 1]]
        nxt = int(np.searchsorted(ts, t, side="right"))
        if nxt == len(ts):
            assert idx == len(ts) - 1 and np.array_equal(pos, P[-1])
        else:
            prev = nxt - 1
            a = (t - ts[prev]) / (ts[nxt] - ts[prev])
            assert idx == prev
            assert np.array_equal(pos, (1.0 - a) * P[prev] + a * P[nxt]), (v, t)
        exact = np.array(_path_at(h, v, t))
        assert np.all(np.abs(pos - exact) <= 4 * np.finfo(float).eps * (1 + np.abs(exact)))
    k = len(queries)
    for verts in stitches:
        n = int(lines[k].split()[1])
        rows = [[float.fromhex(x) for x in l.split()] for l in lines[k + 1:k + 1 + n]]
        k += 1 + n
        T, P = oracle.stitch(h, verts)
        assert n == len(T)
        for row, t, p in zip(rows, T, P):
            assert row[0] == t
            assert np.array_equal(np.array(row[1:]).reshape(-1, 3), p), (verts, t)


REFERENCE_EXAMPLE = "/root/reference/example"


@pytest.mark.skipif(not os.path.isdir(REFERENCE_EXAMPLE), reason="the reference tree is not mounted here")
@pytest.mark.parametrize("driver", ["AlecTest", "testCTCD"])
def test_reference_drivers_compile_unchanged(driver, tmp_path):
    """The reference's own example programs, unchanged (a symbolic link to the file in the reference
    tree, nothing copied), compile against include/ccdamd -- their "../src/*.h" includes resolve to
    include/src -- with the Eigen subset in place of Eigen, and link against libccd_amd.so."""
    lib = os.path.join(ROOT, "collisiondetection_amd", "libccd_amd.so")
    if not os.path.exists(lib):
        pytest.skip("libccd_amd.so is not built")
    ex = tmp_path / "example"
    ex.mkdir()
    src = ex / f"{driver}.cpp"
    os.symlink(os.path.join(REFERENCE_EXAMPLE, f"{driver}.cpp"), src)
    exe = tmp_path / driver
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include", "ccdamd"),
           "-I" + os.path.join(ROOT, "include", "eigen_subset"), str(src), "-o", str(exe),
           "-L" + os.path.dirname(lib), "-lccd_amd", "-Wl,-rpath," + os.path.dirname(lib)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert exe.exists()
