"""The product core under AddressSanitizer + UndefinedBehaviorSanitizer (g++) and MemorySanitizer
(clang), with every runtime index into its per-lane arrays bounds-checked (CCD_CHECK_INDEX), on the
48x48 cloth (two-entry history) and the converging spheres with a three-entry history: the
separating-plane narrowphase with pruned and with full leaf cascades, and the CTCD narrowphase.  The
sanitized builds must run clean, pruned and full cascades must give the same statuses, and the results
must equal the plain host build's.  (A lane-dependent device result was once seen with pruned leaves;
this is the source-level check that the core reads no memory outside its arrays and nothing
uninitialised.)"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from collisiondetection_amd import scenes
from tests.helpers import core_narrowphase, core_separating_plane, three_entry_history

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
BUILD = os.path.join(ROOT, "tests", "_build")
CLANG = "/opt/rocm/llvm/bin/clang++"
FLAGS = ["-O1", "-g", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-DCCD_CHECK_INDEX"]


def _build(kind):
    out = os.path.join(BUILD, f"sanitize_core_{kind}")
    src = os.path.join(NATIVE, "sanitize_core.cpp")
    deps = [src, os.path.join(NATIVE, "core_harness.cpp"),
            os.path.join(ROOT, "collisiondetection_amd", "csrc", "ccd_core.h")]
    if os.path.exists(out) and os.path.getmtime(out) >= max(os.path.getmtime(d) for d in deps):
        return out
    os.makedirs(BUILD, exist_ok=True)
    if kind == "asan":
        cmd = ["g++", *FLAGS, "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
    else:
        cmd = [CLANG, *FLAGS, "-fsanitize=memory", "-fsanitize-memory-track-origins"]
    subprocess.run(cmd + ["-o", out, src], check=True, capture_output=True)
    return out


def _cases():
    s = scenes.cloth(48, drop=0.05, noise=0.01)
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
    s3 = scenes.converging_spheres()
    h3 = oracle.Hist(*three_entry_history(s3))
    vf3, ee3 = oracle.candidates(s3["faces"], h3, s3["eta"], method="tree")
    return [("cloth48_vf", h, vf, False, s["eta"]), ("cloth48_ee", h, ee, True, s["eta"]),
            ("spheres3_vf", h3, vf3, False, s3["eta"]), ("spheres3_ee", h3, ee3, True, s3["eta"])]


def _write(path, h, keys, is_ee, eta):
    with open(path, "wb") as f:
        np.array([h.nverts, len(h.ht), len(keys)], np.int64).tofile(f)
        np.array([int(is_ee)], np.int32).tofile(f)
        np.array([eta], np.float64).tofile(f)
        h.off.astype(np.int64).tofile(f)
        h.ht.astype(np.float64).tofile(f)
        h.hp.astype(np.float64).tofile(f)
        np.ascontiguousarray(keys, np.int32).tofile(f)


@pytest.mark.parametrize("kind", ["asan", "msan"])
def test_core_clean_under_sanitizers(kind, core_harness, tmp_path):
    if kind == "msan" and not os.path.exists(CLANG):
        pytest.skip("clang++ with MemorySanitizer is not installed")
    exe = _build(kind)
    for name, h, keys, is_ee, eta in _cases():
        case, out = tmp_path / f"{name}.bin", tmp_path / f"{name}.out"
        _write(case, h, keys, is_ee, eta)
        r = subprocess.run([exe, str(case), str(out)], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, (name, r.stderr[-3000:])
        assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, (name, r.stderr[-3000:])
        n = len(keys)
        raw = np.fromfile(out, dtype=np.uint8)
        st_p = raw[:4 * n].view(np.int32)
        st_f = raw[4 * n:8 * n].view(np.int32)
        hit = raw[8 * n:12 * n].view(np.int32)
        t = raw[16 * n:24 * n].view(np.float64)
        assert np.array_equal(st_p, st_f), name
        assert np.array_equal(st_p, core_separating_plane(core_harness, h, keys, is_ee, eta)), name
        qs = h.hp.reshape(-1, 3)[h.off[:-1]]
        qe = h.hp.reshape(-1, 3)[h.off[1:] - 1]
        ch, ct, _ = core_narrowphase(core_harness, qs, qe, keys, is_ee, eta)
        assert np.array_equal(hit.astype(bool), ch), name
        assert np.array_equal(t[ch].view(np.int64), ct[ch].view(np.int64)), name
