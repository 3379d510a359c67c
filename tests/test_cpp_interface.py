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
