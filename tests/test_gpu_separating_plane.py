"""GPU separating-plane narrowphase (SeparatingPlaneNarrowPhase.cpp:9-278) against the oracle's recursive
restatement and the host-built core: statuses (1 collision, 0 none, 2 undecided) equal the host core's
on every stencil, with pruned and with full leaf cascades, and the oracle's except where an exact-root
oracle decides a borderline case the device's way."""
import sys

import numpy as np
import pytest

import oracle
from collisiondetection_amd import scenes
from tests.helpers import core_separating_plane, three_entry_history

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def api():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from collisiondetection_amd import api as A

    return A


def _case(kind):
    if kind == "sheets_long_history":
        # the history the velocity filter builds for two colliding sheets: ~9,500 entries per vertex,
        # which the device runs with one wavefront per stencil sharing each trajectory scan
        s = dict(scenes.falling_sheets(3, 0))
        r, _, _, _, h = oracle.velocity_filter(s["qstart"], s["qend"], s["faces"], np.ones(s["qstart"].size),
                                               s["outer_eta"], s["inner_eta"], want_history=True)
        assert r >= 1 and len(h.ht) >= 16 * h.nverts
        return s, h
    if kind == "cloth48_two_entry":
        s = scenes.cloth(48, drop=0.05, noise=0.01)
        return s, oracle.Hist.two_entry(s["qstart"], s["qend"])
    s = scenes.converging_spheres()
    return s, oracle.Hist(*three_entry_history(s))


@pytest.mark.parametrize("kind", ["cloth48_two_entry", "spheres_three_entry", "sheets_long_history"])
def test_gpu_separating_plane_vs_oracle_and_core(api, core_harness, kind):
    say("case: separating plane", kind)
    s, h = _case(kind)
    vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
    r = api.separating_plane_narrowphase(h.off, h.ht, h.hp, vf, ee, s["eta"])
    und = 0
    for keys, is_ee, got in ((vf, False, r["vf_status"]), (ee, True, r["ee_status"])):
        core = core_separating_plane(core_harness, h, keys, is_ee, s["eta"])
        core_harness.core_set_sp_pruning(0)
        try:
            core_full = core_separating_plane(core_harness, h, keys, is_ee, s["eta"])
        finally:
            core_harness.core_set_sp_pruning(1)
        assert np.array_equal(got, core)        # device arithmetic = host arithmetic
        assert np.array_equal(core, core_full)  # pruning changes no status
        oracle.use_rootfinder("aberth")
        try:
            st, margin, depth = oracle.separating_plane(h, keys, s["eta"], is_ee)
        finally:
            oracle.use_default_rootfinder()
        say(kind, "ee" if is_ee else "vf", "gpu", np.bincount(got, minlength=3), "oracle",
            np.bincount(st, minlength=3), "max depth", depth.max())
        diff = np.where(got != st)[0]
        assert all(margin[i] < 1e-12 for i in diff), diff[:5]   # exact roots: only borderline decisions
        if oracle.ref_rpoly_available():
            st_r, margin_r, _ = oracle.separating_plane(h, keys, s["eta"], is_ee)
            diff_r = np.where(got != st_r)[0]
            say("vs the reference's root finder:", len(diff_r), "differences, each decided the device's way by exact roots")
            assert all(margin_r[i] < 1e-12 or st[i] == got[i] for i in diff_r), diff_r[:5]
        und += int((got == 2).sum())
    assert r["undecided"] == und


def test_gpu_separating_plane_class_mirror(api):
    """The reference-interface class: the same sets as the oracle on a two-entry history, and an error
    (instead of the reference's stack overflow) where its recursion does not terminate."""
    say("case: SeparatingPlaneNarrowPhase class")
    s, h = _case("cloth48_two_entry")
    vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
    hist = api.History(s["qstart"])
    hist.finishHistory(s["qend"])
    vfs, ees = set(), set()
    api.SeparatingPlaneNarrowPhase().findCollisions(hist, {(tuple(k), s["eta"]) for k in vf.tolist()},
                                                    {(tuple(k), s["eta"]) for k in ee.tolist()}, vfs, ees)
    for got, keys, is_ee in ((vfs, vf, False), (ees, ee, True)):
        st, _, _ = oracle.separating_plane(h, keys, s["eta"], is_ee)
        assert got == {tuple(k) for k, x in zip(keys.tolist(), st) if x == 1}
    s3, h3 = _case("spheres_three_entry")
    vf3, ee3 = oracle.candidates(s3["faces"], h3, s3["eta"], method="tree")
    st3, _, _ = oracle.separating_plane(h3, ee3, s3["eta"], True)
    assert (st3 == 2).any()
    hist3 = api.History(s3["qstart"])
    for v in range(len(s3["qstart"])):
        e = h3.off[v] + 1
        hist3.addHistory(v, float(h3.ht[e]), h3.hp[3 * e:3 * e + 3])
    hist3.finishHistory(s3["qend"])
    with pytest.raises(RuntimeError, match="does not terminate"):
        api.SeparatingPlaneNarrowPhase().findCollisions(hist3, set(), {(tuple(k), s3["eta"]) for k in ee3.tolist()},
                                                        set(), set())


def test_gpu_separating_plane_empty_and_invalid(api):
    say("case: separating plane edge cases")
    s, h = _case("cloth48_two_entry")
    r = api.separating_plane_narrowphase(h.off, h.ht, h.hp, np.zeros((0, 4), np.int32), np.zeros((0, 4), np.int32),
                                         s["eta"])
    assert len(r["vf_status"]) == 0 and len(r["ee_status"]) == 0 and r["undecided"] == 0
    bad = np.array([[0, 1, 2, h.nverts]], np.int32)
    with pytest.raises(Exception):
        api.separating_plane_narrowphase(h.off, h.ht, h.hp, bad, np.zeros((0, 4), np.int32), s["eta"])


def test_gpu_separating_plane_full_cloth_vs_host_core(api, core_harness):
    """All 7.7M candidate stencils of the 256x256 cloth in one launch per kind, checked against the host
    build of the same code on every stencil: each lane's result must not depend on what the other lanes
    of its wavefront do."""
    say("case: separating plane, full 256^2 cloth vs host core")
    s = scenes.cloth(256)
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    vf, ee = api.kdop_candidates(s["faces"], s["qstart"], s["qend"], s["eta"])
    r = api.separating_plane_narrowphase(h.off, h.ht, h.hp, vf, ee, s["eta"])
    for keys, got, is_ee in ((vf, r["vf_status"], False), (ee, r["ee_status"], True)):
        host = core_separating_plane(core_harness, h, keys, is_ee, s["eta"])
        say("ee" if is_ee else "vf", "statuses", np.bincount(host, minlength=3), "mismatches", int((got != host).sum()))
        assert np.array_equal(got, host)
