"""The 1024^2 crumpled sheet (1,048,576 vertices, 2,093,058 triangles, eta 1e-7) at full size.

The oracle cannot enumerate this mesh's candidates in a test's time, so the full-size run is checked
through properties that do not depend on size:
  * window completeness: for a window W of the grid, a stencil whose four vertices are interior to W
    (every face incident to them lies in W) is generated only by face pairs inside W, so it is a
    full-mesh candidate iff it is a candidate of the submesh W; the oracle enumerates the submesh;
  * no duplicate stencils, every key a valid vertex index;
  * narrowphase: a random sample of the colliding stencils and a random sample of all candidates
    against the oracle (the reference's root finder where built, with exact-root arbitration of every
    disagreement), flags and times of impact.
"""
import numpy as np
import pytest

import oracle
from collisiondetection_amd import scenes
from tests.helpers import compare_results, exact_arbiters

pytestmark = pytest.mark.gpu

N_GRID = 1024


def _void(a):
    return np.ascontiguousarray(a, dtype=np.int32).view(np.dtype((np.void, 16))).ravel()


def _interior_window(n, i0, j0, w):
    """Faces of the vertex window [i0, i0+w) x [j0, j0+w) and the vertices all of whose faces are in it."""
    vi = lambda i, j: n * j + i
    i, j = np.meshgrid(np.arange(i0, i0 + w - 1), np.arange(j0, j0 + w - 1))
    a = vi(i, j).reshape(-1)
    b, c = a + 1, a + n
    d = c + 1
    fw = np.empty((2 * len(a), 3), dtype=np.int32)
    fw[0::2] = np.stack([a, b, d], axis=1)
    fw[1::2] = np.stack([a, d, c], axis=1)
    ii, jj = np.meshgrid(np.arange(i0 + 1, i0 + w - 1), np.arange(j0 + 1, j0 + w - 1))
    interior = np.zeros(n * n, dtype=bool)
    interior[vi(ii, jj).reshape(-1)] = True
    return fw, interior


@pytest.fixture(scope="module")
def sheet():
    return scenes.crumpled_sheet(N_GRID)


def test_gpu_sheet_window_completeness(sheet):
    import torch

    from collisiondetection_amd import api

    s = sheet
    print("case: crumpled sheet 1024^2 candidates", len(s["faces"]), flush=True)
    gvf, gee = api.kdop_candidates(s["faces"], s["qstart"], s["qend"], s["eta"], sort=False, as_numpy=False)
    print("GPU candidates", gvf.shape[0], gee.shape[0], flush=True)
    V = len(s["qstart"])
    for keys in (gvf, gee):
        assert int(keys.min()) >= 0 and int(keys.max()) < V
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    rng = np.random.default_rng(7)
    for trial in range(2):
        w = 48
        i0, j0 = (int(x) for x in rng.integers(0, N_GRID - w, size=2))
        fw, interior = _interior_window(N_GRID, i0, j0, w)
        inside = torch.from_numpy(interior).to(gvf.device)
        ovf, oee = oracle.candidates(fw, h, s["eta"], method="tree")
        for name, g, o in (("vf", gvf, ovf), ("ee", gee, oee)):
            m = inside[g.long()].all(dim=1)
            gsel = g[m].cpu().numpy()
            osel = o[interior[o].all(axis=1)]
            gs, os_ = np.sort(_void(gsel)), np.sort(_void(osel))
            print(f"window ({i0},{j0}) {name}: GPU {len(gs)} oracle {len(os_)}", flush=True)
            assert len(np.unique(gs)) == len(gs)
            assert np.array_equal(gs, os_)
            assert len(os_) > 0
    # no duplicate stencils anywhere: sort a 64-bit mix of each row; equal neighbours are compared in full
    for g in (gvf, gee):
        c = [g[:, k].to(torch.int64) for k in range(4)]
        hsh = ((c[0] * 0x9E3779B1 + c[1]) * 0x85EBCA77 + c[2]) * 0xC2B2AE3D + c[3]
        del c
        hs, order = torch.sort(hsh)
        del hsh
        eq = torch.nonzero(hs[1:] == hs[:-1]).flatten()
        assert len(eq) < 1000, len(eq)   # 64-bit mix collisions are this rare
        for i in eq[:1000].tolist():
            a, b = g[order[i]], g[order[i + 1]]
            assert not bool((a == b).all()), (a.tolist(), b.tolist())
        del hs, order
    del gvf, gee
    torch.cuda.empty_cache()


def test_gpu_sheet_detect_sample(sheet):
    from collisiondetection_amd import api

    s = sheet
    print("case: crumpled sheet 1024^2 detect", flush=True)
    d = api.detect(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
    print("candidates", d["vf_candidates"], d["ee_candidates"], "hits", len(d["vf"]), len(d["ee"]), flush=True)
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    rng = np.random.default_rng(3)
    for pre, is_ee in (("vf", False), ("ee", True)):
        keys, toi = d[pre], d[pre + "_toi"]
        assert len(keys) > 0
        idx = rng.choice(len(keys), size=min(20000, len(keys)), replace=False)
        rec = oracle.narrowphase(h, keys[idx], s["eta"], is_ee)
        rep = compare_results("sheet hits " + pre, np.ones(len(idx), bool), toi[idx], rec,
                              arbiter=exact_arbiters(h, keys[idx], s["eta"], is_ee))
        print(rep, flush=True)
        assert not rep["hard"] and not rep["t_bad_hard"], rep
    # candidates that the oracle finds colliding must be among the GPU's collisions
    gvf, gee = api.kdop_candidates(s["faces"], s["qstart"], s["qend"], s["eta"], sort=False, as_numpy=False)
    for g, pre, is_ee in ((gvf, "vf", False), (gee, "ee", True)):
        import torch

        sel = torch.from_numpy(rng.choice(g.shape[0], size=60000, replace=False)).to(g.device)
        keys = g[sel].cpu().numpy()
        rec = oracle.narrowphase(h, keys, s["eta"], is_ee)
        hits = np.sort(_void(d[pre]))
        q = _void(keys)
        pos = np.minimum(np.searchsorted(hits, q), len(hits) - 1)
        ghit = hits[pos] == q
        rep = compare_results("sheet sample " + pre, ghit, np.zeros(len(keys)), rec, tol_t=np.inf,
                              arbiter=exact_arbiters(h, keys, s["eta"], is_ee))
        print(f"sheet sample {pre}: {int(rec['hit'].sum())} oracle hits of {len(keys)}", rep, flush=True)
        assert not rep["hard"], keys[rep["hard"][:5]]
