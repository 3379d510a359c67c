"""Serving loop (ccd_server_*): each request's result is the same alone and in the mix (bit for bit:
the colliding stencils and their TOIs), refit equals a fresh build, and admission respects the pool."""
import numpy as np
import pytest

from collisiondetection_amd import scenes

pytestmark = pytest.mark.gpu


def _sessions():
    out = []
    for k, s in enumerate(scenes.batched_instances(count=6, seed=21, tri_lo=200, tri_hi=6000)):
        out.append(s)
    out.append(scenes.cloth(64, drop=0.03))
    return out


def _trace(sess, steps, seed=4):
    rng = np.random.default_rng(seed)
    ends = [s["qend"].copy() for s in sess]
    reqs = []
    for r in range(steps):
        for k in rng.permutation(len(sess)):
            ends[k] = ends[k] + rng.normal(0.0, 0.002, size=ends[k].shape)
            reqs.append((int(k), ends[k].copy()))
    return reqs


def _run(sess, reqs, only=None, max_inflight=8, budget=2 << 30):
    from collisiondetection_amd import api

    srv = api.Server(hbm_budget_bytes=budget, max_inflight=max_inflight)
    srv.keep_hits(True)
    ids = [srv.open_session(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"]) for s in sess]
    out = {}
    for rid, (k, q) in enumerate(reqs):
        if only is not None and k != only:
            continue
        srv.submit(ids[k], q, rid)
    srv.drain()
    for r in srv.poll(1 << 16):
        assert r["status"] == 0
        vk, vt, ek, et = srv.fetch_hits(r["request_id"])
        hits = {("vf",) + tuple(x): t for x, t in zip(vk.tolist(), vt)}
        hits.update({("ee",) + tuple(x): t for x, t in zip(ek.tolist(), et)})
        out[r["request_id"]] = (r, hits)
    st = srv.stats()
    srv.close()
    return out, st


def test_server_alone_equals_in_mix():
    print("case: server alone vs in mix", flush=True)
    sess = _sessions()
    reqs = _trace(sess, steps=3)
    mixed, st = _run(sess, reqs)
    assert len(mixed) == len(reqs) and st["pending"] == 0 and st["inflight"] == 0
    for k in range(len(sess)):
        alone, _ = _run(sess, reqs, only=k, max_inflight=1)
        for rid, (r, hits) in alone.items():
            rm, hm = mixed[rid]
            assert hits == hm, (k, rid)
            assert (r["checksum_vf"], r["checksum_ee"]) == (rm["checksum_vf"], rm["checksum_ee"])
            assert (r["vf_candidates"], r["ee_candidates"]) == (rm["vf_candidates"], rm["ee_candidates"])


def test_server_refit_equals_fresh_detect():
    from collisiondetection_amd import api

    print("case: server refit vs fresh build", flush=True)
    sess = _sessions()[:3] + [_sessions()[-1]]
    reqs = _trace(sess, steps=2, seed=9)
    res, _ = _run(sess, reqs)
    prev = [s["qend"].copy() for s in sess]
    for rid, (k, q) in enumerate(reqs):
        s = sess[k]
        d = api.detect(s["faces"], prev[k], q, s["eta"], s["eta"])
        r, hits = res[rid]
        assert (r["vf_candidates"], r["ee_candidates"]) == (d["vf_candidates"], d["ee_candidates"])
        fresh = {("vf",) + tuple(x): t for x, t in zip(d["vf"].tolist(), d["vf_toi"])}
        fresh.update({("ee",) + tuple(x): t for x, t in zip(d["ee"].tolist(), d["ee_toi"])})
        assert hits == fresh, rid
        prev[k] = q


def test_server_small_pool_admission():
    print("case: server admission with a small pool", flush=True)
    sess = _sessions()
    reqs = _trace(sess, steps=2, seed=5)
    big, st_big = _run(sess, reqs)
    # a pool that holds about one of the largest requests at a time makes requests wait for memory
    budget = st_big["max_request_bytes"] * 3 // 2 + 4096
    small, st = _run(sess, reqs, budget=budget)
    print("small pool", budget, st, flush=True)
    assert st["completed"] == len(reqs)
    for rid in big:
        assert big[rid][1] == small[rid][1]


def test_server_overflow_reruns():
    """A request whose stencils outgrow the workspace sized from the session's last counts (a nearly
    static request, then a large motion): the narrowphase reads only the stored stencils, the request
    is rerun with a larger workspace, and its result equals a fresh detection."""
    from collisiondetection_amd import api

    print("case: server workspace overflow and rerun", flush=True)
    s = scenes.cloth(48, drop=0.0, noise=0.0)   # qend == qstart: a static sheet
    rng = np.random.default_rng(2)
    q1 = s["qstart"] + rng.normal(0.0, 1e-5, size=s["qstart"].shape)   # sizes the next workspace small
    q = q1 + np.array([0.0, 0.0, -0.05]) + rng.normal(0.0, 0.02, size=s["qstart"].shape)
    srv = api.Server(hbm_budget_bytes=1 << 30, max_inflight=2)
    srv.keep_hits(True)
    sid = srv.open_session(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"])
    srv.submit(sid, q1, 0)
    srv.drain()
    srv.poll()
    reruns0 = srv.stats()["reruns"]
    srv.submit(sid, q, 1)
    srv.drain()
    (r,) = srv.poll()
    st = srv.stats()
    vk, vt, ek, et = srv.fetch_hits(1)
    srv.close()
    print("stats", st, "result", r, flush=True)
    assert r["status"] == 0 and st["reruns"] > reruns0
    d = api.detect(s["faces"], q1, q, s["eta"], s["eta"])
    assert (r["vf_candidates"], r["ee_candidates"]) == (d["vf_candidates"], d["ee_candidates"])
    assert {tuple(k): t for k, t in zip(vk.tolist(), vt)} == {tuple(k): t for k, t in zip(d["vf"].tolist(), d["vf_toi"])}
    assert {tuple(k): t for k, t in zip(ek.tolist(), et)} == {tuple(k): t for k, t in zip(d["ee"].tolist(), d["ee_toi"])}


def test_server_full_workload_alone_equals_in_mix():
    """The served workload's 256 resident sessions (seed 4, 1k-500k triangles, 20.6M resident): a
    sampled trace of requests (each a fresh N(0, 0.002^2) step of the session's end positions) runs
    in the mix with 16 requests in flight; for a sample of sessions, the same requests run alone on a
    server holding only that session, and through fresh detections (refit = rebuild): the colliding
    stencils and time-of-impact bits are the same in all three."""
    from collisiondetection_amd import api

    print("case: served workload, 256 sessions", flush=True)
    sess = scenes.served_sessions(256)
    rng = np.random.default_rng(4)
    ends = [s["qend"].copy() for s in sess]
    reqs = []
    for _ in range(400):
        k = int(rng.integers(0, len(sess)))
        ends[k] = ends[k] + rng.normal(0.0, 0.002, size=ends[k].shape)
        reqs.append((k, ends[k].copy()))
    srv = api.Server(hbm_budget_bytes=32 << 30, max_inflight=16)
    srv.keep_hits(True)
    ids = []
    for j, s in enumerate(sess):
        ids.append(srv.open_session(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"]))
        if j % 64 == 63:
            print("opened", j + 1, "sessions", flush=True)
    for rid, (k, q) in enumerate(reqs):
        srv.submit(ids[k], q, rid)
    srv.drain()
    mixed = {}
    for r in srv.poll(1 << 16):
        assert r["status"] == 0, r
        vk, vt, ek, et = srv.fetch_hits(r["request_id"])
        hits = {("vf",) + tuple(x): t for x, t in zip(vk.tolist(), vt)}
        hits.update({("ee",) + tuple(x): t for x, t in zip(ek.tolist(), et)})
        mixed[r["request_id"]] = (r, hits)
    st = srv.stats()
    srv.close()
    print("mixed run", st, flush=True)
    assert len(mixed) == len(reqs)
    nf = np.array([len(s["faces"]) for s in sess])
    used = sorted({k for k, _ in reqs})
    big = [k for k in used if nf[k] > 100_000]
    sample = list(rng.choice([k for k in used if nf[k] <= 100_000], size=5, replace=False)) + big[:1]
    for k in sample:
        k = int(k)
        mine = [(rid, q) for rid, (kk, q) in enumerate(reqs) if kk == k]
        alone, _ = _run([sess[k]], [(0, q) for _, q in mine], max_inflight=1)
        prev = sess[k]["qend"].copy()
        for j, (rid, q) in enumerate(mine):
            ra, ha = alone[j]
            rm, hm = mixed[rid]
            assert ha == hm, (k, rid)
            assert (ra["vf_candidates"], ra["ee_candidates"]) == (rm["vf_candidates"], rm["ee_candidates"])
            d = api.detect(sess[k]["faces"], prev, q, sess[k]["eta"], sess[k]["eta"])
            fresh = {("vf",) + tuple(x): t for x, t in zip(d["vf"].tolist(), d["vf_toi"])}
            fresh.update({("ee",) + tuple(x): t for x, t in zip(d["ee"].tolist(), d["ee_toi"])})
            assert fresh == hm, (k, rid)
            prev = q
        print("session", k, nf[k], "triangles:", len(mine), "requests equal alone, in the mix and fresh", flush=True)
