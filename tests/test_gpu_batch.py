"""Batched detection (ccd_batch_detect): many independent instances in one pipeline equal the union of
per-instance ccd_detect calls and the oracle; the captured HIP graph replays give the same results,
also after the positions behind the captured pointers change."""
import numpy as np
import pytest

import oracle
from collisiondetection_amd import scenes
from tests.helpers import compare_results, exact_arbiters

pytestmark = pytest.mark.gpu


def _as_set(keys, toi):
    return {tuple(k): t for k, t in zip(keys.tolist(), toi)}


def _instances(count=24, seed=7):
    return scenes.batched_instances(count=count, seed=seed, tri_lo=80, tri_hi=1500)


def test_batch_equals_per_instance_detect():
    from collisiondetection_amd import api

    print("case: batch vs per-instance", flush=True)
    inst = _instances()
    b = api.Batch(inst)
    r = b.detect(1e-6, 1e-6)
    per_vf, per_ee, nvf, nee = {}, {}, 0, 0
    for k, s in enumerate(inst):
        d = api.detect(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
        off = b.vert_off[k]
        per_vf.update(_as_set(d["vf"] + off, d["vf_toi"]))
        per_ee.update(_as_set(d["ee"] + off, d["ee_toi"]))
        nvf += d["vf_candidates"]
        nee += d["ee_candidates"]
    assert r["vf_candidates"] == nvf and r["ee_candidates"] == nee
    gv, ge = _as_set(r["vf"], r["vf_toi"]), _as_set(r["ee"], r["ee_toi"])
    assert gv == per_vf and ge == per_ee   # same stencils, same TOI bits
    assert len(gv) + len(ge) > 0
    # no stencil spans two instances
    for keys in (r["vf"], r["ee"]):
        inst_ids = b.instance_of_vertex(keys)
        assert np.all(inst_ids == inst_ids[:, :1])


def test_batch_instances_vs_oracle():
    from collisiondetection_amd import api

    print("case: batch vs oracle", flush=True)
    inst = _instances(count=6, seed=11)
    b = api.Batch(inst)
    r = b.detect(1e-6, 1e-6)
    gv, ge = _as_set(r["vf"], r["vf_toi"]), _as_set(r["ee"], r["ee_toi"])
    seen = 0
    for k, s in enumerate(inst):
        h = oracle.Hist.two_entry(s["qstart"], s["qend"])
        vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
        off = int(b.vert_off[k])
        for keys, is_ee, got in ((vf, False, gv), (ee, True, ge)):
            glob = [tuple(int(v) + off for v in x) for x in keys.tolist()]
            hit = np.array([g in got for g in glob], dtype=bool)
            t = np.array([got.get(g, 0.0) for g in glob])
            seen += int(hit.sum())
            rec = oracle.narrowphase(h, keys, s["eta"], is_ee)
            rep = compare_results(f"inst{k}" + ("ee" if is_ee else "vf"), hit, t, rec,
                                  arbiter=exact_arbiters(h, keys, s["eta"], is_ee))
            print(rep, flush=True)
            assert not rep["hard"] and not rep["t_bad_hard"], rep
    # every batched hit belongs to some instance's oracle candidate set
    assert seen == len(gv) + len(ge)


def test_batch_graph_replay_and_refit():
    import torch

    from collisiondetection_amd import api

    print("case: graph capture and replay", flush=True)
    inst = _instances(count=16, seed=13)
    b = api.Batch(inst)
    r0 = b.detect(1e-6, 1e-6, graph=False)
    r1 = b.detect(1e-6, 1e-6, graph=True)    # synchronous run + capture
    r2 = b.detect(1e-6, 1e-6, graph=True)    # replay
    for r in (r1, r2):
        assert _as_set(r["vf"], r["vf_toi"]) == _as_set(r0["vf"], r0["vf_toi"])
        assert _as_set(r["ee"], r["ee_toi"]) == _as_set(r0["ee"], r0["ee_toi"])
    # new end positions behind the same pointers: the replay must equal a fresh synchronous run
    rng = np.random.default_rng(3)
    for step in range(3):
        noise = torch.from_numpy(rng.normal(0, 0.002, size=b.V * 3)).to(b.qe.device)
        b.qe.add_(noise)
        torch.cuda.synchronize()
        rg = b.detect(1e-6, 1e-6, graph=True)
        rs = b.detect(1e-6, 1e-6, graph=False)
        assert _as_set(rg["vf"], rg["vf_toi"]) == _as_set(rs["vf"], rs["vf_toi"]), step
        assert _as_set(rg["ee"], rg["ee_toi"]) == _as_set(rs["ee"], rs["ee_toi"]), step
        assert rg["vf_candidates"] == rs["vf_candidates"] and rg["ee_candidates"] == rs["ee_candidates"]


def test_autotune_keeps_results():
    from collisiondetection_amd import api

    print("case: autotune", flush=True)
    s = scenes.cloth(96, drop=0.05)
    d0 = api.detect(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
    shape = api.autotune(s["qstart"], s["qend"], s["eta"])
    assert len(shape) == 8 and shape[0] in (64, 128, 256) and shape[3] in (64, 128, 256)
    assert shape[6] in (0, 1) and shape[7] in (0, 1)
    d1 = api.detect(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
    assert _as_set(d0["vf"], d0["vf_toi"]) == _as_set(d1["vf"], d1["vf_toi"])
    assert _as_set(d0["ee"], d0["ee_toi"]) == _as_set(d1["ee"], d1["ee_toi"])
    # the grouping switched either way leaves every result as it was
    for g in (0, 1):
        api.set_queue_shape(shape[:6] + [g, 1 - g])
        d2 = api.detect(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
        assert _as_set(d0["vf"], d0["vf_toi"]) == _as_set(d2["vf"], d2["vf_toi"]), g
        assert _as_set(d0["ee"], d0["ee_toi"]) == _as_set(d2["ee"], d2["ee_toi"]), g
    api.set_queue_shape(shape)


def test_edge_edge_split_keeps_results():
    """The two-pass edge-edge narrowphase (the default) and the one-pass form give the same colliding
    stencils and time-of-impact bits: in a batch (synchronous run and graph replays) and for one mesh."""
    from collisiondetection_amd import api

    print("case: edge-edge split on / off", flush=True)
    inst = scenes.batched_instances(count=16, seed=5, tri_lo=400, tri_hi=4000)
    b = api.Batch(inst)
    out = {}
    try:
        for on in (True, False):
            api.set_edge_edge_split(on, ctx=b.ctx)
            rs = [b.detect(1e-6, 1e-6, graph=g) for g in (False, True, True)]
            for r in rs[1:]:
                assert _as_set(r["ee"], r["ee_toi"]) == _as_set(rs[0]["ee"], rs[0]["ee_toi"]), on
                assert _as_set(r["vf"], r["vf_toi"]) == _as_set(rs[0]["vf"], rs[0]["vf_toi"]), on
            out[on] = rs[0]
        print("ee candidates", out[True]["ee_candidates"], "hits", len(out[True]["ee"]), flush=True)
        assert _as_set(out[True]["ee"], out[True]["ee_toi"]) == _as_set(out[False]["ee"], out[False]["ee_toi"])
        assert _as_set(out[True]["vf"], out[True]["vf_toi"]) == _as_set(out[False]["vf"], out[False]["vf_toi"])
        assert len(out[True]["ee"]) > 0
        s = scenes.converging_spheres()
        d = {}
        for on in (True, False):
            api.set_edge_edge_split(on)
            d[on] = api.detect(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
        assert _as_set(d[True]["ee"], d[True]["ee_toi"]) == _as_set(d[False]["ee"], d[False]["ee_toi"])
        assert len(d[True]["ee"]) > 0
    finally:
        api.set_edge_edge_split(True, ctx=b.ctx)
        api.set_edge_edge_split(True)


def test_batch_graph_after_buffer_regrow():
    """A captured pipeline holds raw buffer addresses.  Detections of other meshes on the same context
    that grow buffers sized by vertex or face count (without changing the stencil capacities) must make
    the next graph call recapture, not replay into freed memory: its result equals a synchronous run."""
    from collisiondetection_amd import api

    print("case: graph after buffer regrowth", flush=True)
    inst = _instances(count=8, seed=17)
    b = api.Batch(inst)
    r0 = b.detect(1e-6, 1e-6, graph=False)
    b.detect(1e-6, 1e-6, graph=True)   # synchronous run + capture
    # many vertices, few faces: vertex-sized buffers grow, pair and stencil counts stay small
    rng = np.random.default_rng(5)
    V = 8 * b.V
    q = rng.uniform(-50, 50, size=(V, 3))
    faces = np.arange(30, dtype=np.int32).reshape(10, 3)
    api.detect(faces, q, q, 1e-6, 1e-6)
    rg = b.detect(1e-6, 1e-6, graph=True)
    assert _as_set(rg["vf"], rg["vf_toi"]) == _as_set(r0["vf"], r0["vf_toi"])
    assert _as_set(rg["ee"], rg["ee_toi"]) == _as_set(r0["ee"], r0["ee_toi"])
    # more faces, spread out so that they pair with nothing: face-sized buffers grow
    b.detect(1e-6, 1e-6, graph=True)
    g = scenes.cloth(int(np.sqrt(b.F)) + 8)
    flat = np.array(g["qstart"], dtype=np.float64)
    flat[:, 2] = 0.0
    api.detect(g["faces"], flat, flat, 1e-9, 1e-9)
    rg = b.detect(1e-6, 1e-6, graph=True)
    assert _as_set(rg["vf"], rg["vf_toi"]) == _as_set(r0["vf"], r0["vf_toi"])
    assert _as_set(rg["ee"], rg["ee_toi"]) == _as_set(r0["ee"], r0["ee_toi"])


def test_batch_full_workload():
    """The 512-instance batched workload itself (seed 2; 80-20,480 triangles per mesh, 4.0M triangles):
    the batched pipeline, its captured-graph replay and the per-instance (plain) detections return the
    same colliding stencils with the same time-of-impact bits; a sample of instances equals the oracle."""
    from collisiondetection_amd import api

    print("case: batched workload, 512 instances", flush=True)
    inst = scenes.batched_instances()
    b = api.Batch(inst)
    print("instances", b.ninst, "triangles", b.F, flush=True)
    r0 = b.detect(1e-6, 1e-6, graph=False)
    b.detect(1e-6, 1e-6, graph=True)          # synchronous run + capture
    rg = b.detect(1e-6, 1e-6, graph=True)     # replay
    ref_vf, ref_ee = _as_set(r0["vf"], r0["vf_toi"]), _as_set(r0["ee"], r0["ee_toi"])
    assert _as_set(rg["vf"], rg["vf_toi"]) == ref_vf and _as_set(rg["ee"], rg["ee_toi"]) == ref_ee
    assert (rg["vf_candidates"], rg["ee_candidates"]) == (r0["vf_candidates"], r0["ee_candidates"])
    print("candidates", r0["vf_candidates"], r0["ee_candidates"], "collisions", len(ref_vf), len(ref_ee), flush=True)
    per_vf, per_ee, nvf, nee = {}, {}, 0, 0
    for k, s in enumerate(inst):
        d = api.detect(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
        off = b.vert_off[k]
        per_vf.update(_as_set(d["vf"] + off, d["vf_toi"]))
        per_ee.update(_as_set(d["ee"] + off, d["ee_toi"]))
        nvf += d["vf_candidates"]
        nee += d["ee_candidates"]
    assert (nvf, nee) == (r0["vf_candidates"], r0["ee_candidates"])
    assert per_vf == ref_vf and per_ee == ref_ee
    # oracle on a sample of the instances (the smaller half, where the oracle is quick)
    rng = np.random.default_rng(2)
    small = [k for k, s in enumerate(inst) if len(s["faces"]) <= 3000]
    for k in rng.choice(small, size=6, replace=False):
        s = inst[int(k)]
        h = oracle.Hist.two_entry(s["qstart"], s["qend"])
        vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
        off = int(b.vert_off[int(k)])
        for keys, is_ee, got in ((vf, False, ref_vf), (ee, True, ref_ee)):
            glob = [tuple(int(v) + off for v in x) for x in keys.tolist()]
            hit = np.array([g in got for g in glob], dtype=bool)
            t = np.array([got.get(g, 0.0) for g in glob])
            rec = oracle.narrowphase(h, keys, s["eta"], is_ee)
            rep = compare_results(f"inst{k}" + ("ee" if is_ee else "vf"), hit, t, rec,
                                  arbiter=exact_arbiters(h, keys, s["eta"], is_ee))
            print(rep["name"], len(keys), "stencils", rep["hits"], "hits", rep["n_ref_root_error"], "ref root errors",
                  flush=True)
            assert not rep["hard"] and not rep["t_bad_hard"], rep
