"""Benchmark: swept triangles per second of the full swept-primitive CCD path on one MI355X.

Workload (default): the draped cloth self-CCD step -- a 256 x 256 vertex grid (65,536 vertices,
130,050 triangles) folded as z = 0.15 sin(6x) cos(5y), end = start - 0.02 z + N(0, 0.004^2), eta = 1e-6,
FP64.  One step = the reference's KDOPBroadPhase (leaf bounds, tree, traversal, stencil set) followed by
CTCDNarrowPhase on every candidate stencil, with positions already resident in HBM; the colliding
stencils and their times of impact are the output.

Printed: one JSON line with the metric, the per-stage HIP-event timings, the roofline of the dominant
kernel and a CPU baseline (the oracle's reference-structured CPU path on a bounded sample), plus
secondary legs: the isolating root mode, the separating-plane narrowphase on the same candidates, the
velocity filter (the reference's example input and a falling-sheets contact problem), batched
instances and the serving loop (each can be switched off with its --no-* flag).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, sample_stencils=60000, seed=0):
    """The oracle's reference-structured CPU path (top-down k-DOP tree, recursive traversal with std::set
    insertion, then the per-stencil narrowphase), single-threaded, on a bounded sample: the broadphase
    in full and the narrowphase on a random sample of its stencils, extrapolated per stencil."""
    import numpy as np

    import oracle

    rf = oracle.use_default_rootfinder()
    h = oracle.Hist.two_entry(scene["qstart"], scene["qend"])
    t0 = time.perf_counter()
    vf, ee = oracle.candidates(scene["faces"], h, scene["eta"], method="tree_set")
    t_broad = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    nv = min(len(vf), sample_stencils * len(vf) // max(1, len(vf) + len(ee)))
    ne = min(len(ee), sample_stencils - nv)
    sv = vf[rng.choice(len(vf), size=nv, replace=False)] if nv else vf[:0]
    se = ee[rng.choice(len(ee), size=ne, replace=False)] if ne else ee[:0]
    t0 = time.perf_counter()
    oracle.narrowphase(h, sv, scene["eta"], False)
    t_vf = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.narrowphase(h, se, scene["eta"], True)
    t_ee = time.perf_counter() - t0
    t_narrow = (t_vf / max(nv, 1)) * len(vf) + (t_ee / max(ne, 1)) * len(ee)
    total = t_broad + t_narrow
    F = len(scene["faces"])
    return {"value": F / total, "unit": "swept tris/s", "cores": 1, "kind": "port",
            "sample": (f"oracle with root finder '{rf}'; full broadphase ({t_broad:.2f} s, {len(vf)} VF + {len(ee)} EE stencils) + narrowphase on "
                       f"{nv} VF and {ne} EE sampled stencils ({t_vf + t_ee:.2f} s), extrapolated to all: "
                       f"{total:.1f} s per step")}


def bench_separating_plane(ctx, scene, qs, qe, steps, log, cpu_sample=30000, seed=1):
    """SeparatingPlaneNarrowPhase (the narrowphase VelocityFilter uses) on the workload's candidate
    stencils with a two-entry history built on the device, reference-root mode: stencils per second;
    the oracle's recursive restatement timed on a random sample of the same stencils (1 thread)."""
    import ctypes

    import numpy as np
    import torch

    import oracle
    from collisiondetection_amd import _native as N
    from collisiondetection_amd import api

    L = N.lib()
    dev = qs.device
    V = qs.numel() // 3
    off = torch.arange(0, 2 * V + 1, 2, dtype=torch.int64, device=dev)
    ht = torch.tensor([0.0, 1.0], dtype=torch.float64, device=dev).repeat(V)
    hp = torch.stack([qs.reshape(-1, 3), qe.reshape(-1, 3)], dim=1).reshape(-1).contiguous()
    vk, ek = api.kdop_candidates(scene["faces"], scene["qstart"], scene["qend"], scene["eta"], ctx=ctx,
                                 as_numpy=False)
    vk, ek = vk.reshape(-1).contiguous(), ek.reshape(-1).contiguous()
    nv, ne = vk.numel() // 4, ek.numel() // 4
    vs = torch.empty(max(nv, 1), dtype=torch.uint8, device=dev)
    es = torch.empty(max(ne, 1), dtype=torch.uint8, device=dev)
    und = ctypes.c_int64(0)
    eta = float(scene["eta"])

    def run():
        ctx.check(L.ccd_separating_plane_narrowphase(ctx.ptr, off.data_ptr(), ht.data_ptr(), hp.data_ptr(), V, 2 * V,
                                                     vk.data_ptr(), None, nv, ek.data_ptr(), None, ne, eta,
                                                     vs.data_ptr(), es.data_ptr(), ctypes.byref(und)))

    run()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize(dev)
    wall = (time.perf_counter() - t0) / steps
    n = nv + ne
    log(f"separating plane: {n} stencils in {wall * 1e3:.2f} ms")
    out = {"stencils": n, "ms": wall * 1e3, "stencils_per_s": n / wall,
           "status": {"vf": np.bincount(vs[:nv].cpu().numpy(), minlength=3).tolist(),
                      "ee": np.bincount(es[:ne].cpu().numpy(), minlength=3).tolist()},
           "undecided": und.value}
    # CPU: the oracle (recursive, as the reference) on a random sample of the same stencils
    rf = oracle.use_default_rootfinder()
    h = oracle.Hist.two_entry(scene["qstart"], scene["qend"])
    rng = np.random.default_rng(seed)
    kv = vk.reshape(-1, 4).cpu().numpy()
    ke = ek.reshape(-1, 4).cpu().numpy()
    sv = kv[rng.choice(nv, size=min(nv, cpu_sample * nv // max(n, 1)), replace=False)] if nv else kv
    se = ke[rng.choice(ne, size=min(ne, cpu_sample - len(sv)), replace=False)] if ne else ke
    t0 = time.perf_counter()
    oracle.separating_plane(h, sv, eta, False)
    oracle.separating_plane(h, se, eta, True)
    tc = time.perf_counter() - t0
    per = tc / max(len(sv) + len(se), 1)
    out["cpu_baseline"] = {"value": 1.0 / per, "unit": "stencils/s", "cores": 1, "kind": "port",
                           "sample": f"oracle (root finder '{rf}') on {len(sv)} VF + {len(se)} EE random stencils "
                                     f"of the same set, {tc:.2f} s"}
    return out


def bench_velocity_filter(ctx, log):
    """VelocityFilter::velocityFilter: the reference's own example (example/mesh1.obj -> mesh2.obj, kept as
    tests/golden/velocity_filter_{start,end}.obj, testVelocityFilter's parameters) and a falling-sheets
    contact problem (5x5 sheets, lower one fixed), seconds per call; the oracle (the reference's shape,
    1 thread) timed on the sheets."""
    import numpy as np

    import oracle
    from collisiondetection_amd import api, scenes

    def load_obj(path):
        v, f = [], []
        with open(path) as fh:
            for line in fh:
                t = line.split()
                if t and t[0] == "v":
                    v.extend(float(x) for x in t[1:4])
                elif t and t[0] == "f":
                    f.extend(int(x) - 1 for x in t[1:4])
        return np.array(v), np.array(f, dtype=np.int32)

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    q1, f = load_obj(os.path.join(here, "velocity_filter_start.obj"))
    q2, _ = load_obj(os.path.join(here, "velocity_filter_end.obj"))
    s = scenes.falling_sheets(5, 3)
    im = np.ones(s["qstart"].size)
    im[:75] = 0.0
    cases = {"reference_example": (q1, q2, f, np.ones_like(q1), 2e-8, 1e-8),
             "falling_sheets5": (s["qstart"], s["qend"], s["faces"], im, s["outer_eta"], s["inner_eta"])}
    out = {}
    for name, (a, b, faces, masses, outer, inner) in cases.items():
        t0 = time.perf_counter()
        r, _, st = api.velocity_filter(a, b, faces, masses, outer, inner, ctx=ctx)
        wall = time.perf_counter() - t0
        log(f"velocity filter {name}: result {r}, {st['firings']} firings, {wall:.2f} s "
            f"(host loop {st['host_seconds']:.2f} s, device {st['device_seconds']:.3f} s)")
        out[name] = {"result": r, "seconds": wall, "host_seconds": st["host_seconds"],
                     "device_seconds": st["device_seconds"], "iterations": st["iterations"],
                     "layers": st["layers"], "firings": st["firings"]}
    rf = oracle.use_default_rootfinder()
    a, b, faces, masses, outer, inner = cases["falling_sheets5"]
    t0 = time.perf_counter()
    oracle.velocity_filter(a, b, faces, masses, outer, inner)
    tc = time.perf_counter() - t0
    out["falling_sheets5"]["cpu_baseline"] = {"value": tc, "unit": "s per call", "cores": 1, "kind": "port",
                                              "sample": f"oracle (root finder '{rf}'), the whole call"}
    return out


def bench_batched(steps, log):
    """The 512-instance mixed workload: plain (ccd_detect per instance, the same kernels launched per
    instance) vs batched (one ccd_batch_detect pipeline) vs batched replayed from a captured HIP graph."""
    import ctypes

    import numpy as np
    import torch

    from collisiondetection_amd import _native as N
    from collisiondetection_amd import api, scenes

    inst = scenes.batched_instances()
    b = api.Batch(inst)
    ctx = b.ctx
    L = N.lib()
    F = b.F
    log(f"batched workload: {b.ninst} instances, {F} triangles, {b.V} vertices")
    # per-instance face arrays with local vertex indices (the plain path's input)
    local_faces = [api.to_device(s["faces"], torch.int32, ctx).reshape(-1) for s in inst]
    cap = max(1024, 2 * max(len(s["faces"]) for s in inst))
    dev = ctx.torch_device
    vk = torch.empty(4 * cap, dtype=torch.int32, device=dev)
    vt = torch.empty(cap, dtype=torch.float64, device=dev)
    ek = torch.empty(4 * cap, dtype=torch.int32, device=dev)
    et = torch.empty(cap, dtype=torch.float64, device=dev)
    c = [ctypes.c_int64(0) for _ in range(4)]
    qs0, qe0 = b.qs.data_ptr(), b.qe.data_ptr()

    def plain():
        hits = 0
        for k, s in enumerate(inst):
            off = int(b.vert_off[k]) * 24
            rc = L.ccd_detect(ctx.ptr, local_faces[k].data_ptr(), len(s["faces"]), qs0 + off, qe0 + off,
                              len(s["qstart"]), s["eta"], s["eta"], None, vk.data_ptr(), vt.data_ptr(), cap,
                              ctypes.byref(c[0]), ek.data_ptr(), et.data_ptr(), cap, ctypes.byref(c[1]),
                              ctypes.byref(c[2]), ctypes.byref(c[3]))
            ctx.check(rc)
            hits += c[0].value + c[1].value
        return hits

    ctx.bind_stream()
    plain()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    np_ = max(1, steps // 4)
    for _ in range(np_):
        h_plain = plain()
    torch.cuda.synchronize(dev)
    t_plain = (time.perf_counter() - t0) / np_
    r = b.detect(1e-6, 1e-6)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = b.detect(1e-6, 1e-6)
    torch.cuda.synchronize(dev)
    t_batch = (time.perf_counter() - t0) / steps
    h_batch = len(r["vf"]) + len(r["ee"])
    b.detect(1e-6, 1e-6, graph=True)   # synchronous run + capture
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        rg = b.detect(1e-6, 1e-6, graph=True)
    torch.cuda.synchronize(dev)
    t_graph = (time.perf_counter() - t0) / steps
    h_graph = len(rg["vf"]) + len(rg["ee"])
    log(f"batched: plain {t_plain * 1e3:.2f} ms, batched {t_batch * 1e3:.2f} ms, graph {t_graph * 1e3:.2f} ms; "
        f"collisions {h_plain} / {h_batch} / {h_graph}")
    return {"instances": b.ninst, "triangles": F, "plain_tris_per_s": F / t_plain, "batched_tris_per_s": F / t_batch,
            "graph_tris_per_s": F / t_graph, "speedup_vs_plain": t_plain / t_graph,
            "ms": {"plain": t_plain * 1e3, "batched": t_batch * 1e3, "graph": t_graph * 1e3},
            "collisions_equal": h_plain == h_batch == h_graph, "collisions": h_graph,
            "candidates": {"vf": rg["vf_candidates"], "ee": rg["ee_candidates"]}}


def bench_served(seconds, rates, log, sessions=256, budget_gib=32, max_inflight=16):
    """The served workload: `sessions` device-resident sessions (log-uniform 1k-500k triangles, seed 4);
    for each rate, Poisson arrivals at that many requests/s for `seconds`, each to a uniform session,
    supplying new end positions = previous end + N(0, 0.002^2) (the sessions carry on from run to run).
    Latency = request submission to result on the host.  A run stops offering load once two seconds of
    requests are waiting (overload), then drains."""
    import numpy as np

    from collisiondetection_amd import api, scenes

    sess = scenes.served_sessions(sessions)
    nf = np.array([len(s["faces"]) for s in sess])
    log(f"served workload: {len(sess)} sessions, {int(nf.sum())} resident triangles")
    srv = api.Server(hbm_budget_bytes=budget_gib << 30, max_inflight=max_inflight)
    t0 = time.perf_counter()
    ids = []
    for j, s in enumerate(sess):
        ids.append(srv.open_session(s["faces"], s["qstart"], s["qend"], s["eta"], s["eta"]))
        if j % 64 == 63:
            log(f"opened {j + 1} sessions")
    log(f"sessions opened in {time.perf_counter() - t0:.1f} s; stats {srv.stats()}")
    ends = [s["qend"].reshape(-1).copy() for s in sess]
    noise = scenes.NoiseBank(0.002, seed=4)
    rng = np.random.default_rng(4)
    runs = []
    rid = 0
    for rate in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 1.2) + 16))
        arrivals = arrivals[arrivals < seconds]
        targets = rng.integers(0, len(sess), size=len(arrivals))
        results = []
        submitted = 0
        backlog_cap = max(16, int(rate * 2))
        overloaded = False
        st0 = srv.stats()
        start = time.perf_counter()
        for i, (ta, k) in enumerate(zip(arrivals, targets)):
            while True:
                now = time.perf_counter() - start
                if now >= ta:
                    break
                results += srv.poll()
                if ta - (time.perf_counter() - start) > 0.0005:
                    time.sleep(0.0002)
            ends[k] = ends[k] + noise.draw(len(ends[k]))
            srv.submit(ids[k], ends[k], rid)
            rid += 1
            submitted += 1
            if i % 64 == 0:
                st = srv.stats()
                if st["pending"] > backlog_cap:
                    overloaded = True
                    log(f"rate {rate}: overloaded at request {i} ({ta:.2f} s): {st['pending']} pending")
                    break
            if i % 2000 == 0:
                log(f"rate {rate}: {i} requests submitted at {ta:.1f} s, {len(results)} done")
        offered_s = time.perf_counter() - start
        srv.drain()
        results += srv.poll(1 << 20)
        end = time.perf_counter() - start
        st = srv.stats()
        ok = [r for r in results if r["status"] == 0]
        lat = np.array([r["t_done"] - r["t_submit"] for r in ok]) * 1e3
        tris = np.array([nf[r["session"]] for r in ok])
        small = lat[tris <= 100000]
        t_first = min(r["t_submit"] for r in ok)
        t_last = max(r["t_done"] for r in ok)
        out = {"rate_offered": rate, "seconds": seconds, "submitted": submitted, "completed": len(ok),
               "failed": len(results) - len(ok), "overloaded": overloaded, "offered_wall_s": offered_s,
               "drain_wall_s": end - offered_s, "swept_tris_per_s": float(tris.sum() / (t_last - t_first)),
               "requests_per_s": len(ok) / (t_last - t_first),
               "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                              "p50_le100k": float(np.percentile(small, 50)) if len(small) else None,
                              "p99_le100k": float(np.percentile(small, 99)) if len(small) else None},
               "reruns": st["reruns"] - st0["reruns"], "mem_waits": st["mem_waits"] - st0["mem_waits"]}
        log("served:", json.dumps(out))
        runs.append(out)
    srv.close()
    return {"sessions": len(sess), "resident_triangles": int(nf.sum()), "budget_gib": budget_gib,
            "max_inflight": max_inflight, "runs": runs,
            "client_noise": "one bank of 2^24 N(0, 0.002^2) draws, read at a random offset per request"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=256, help="cloth grid size (256 = the 130,050-triangle cloth)")
    ap.add_argument("--workload", choices=["cloth", "sheet"], default="cloth",
                    help="cloth: the draped cloth (default); sheet: the crumpled sheet (--grid 1024 = 2,093,058 tris)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-separating-plane", action="store_true")
    ap.add_argument("--no-velocity-filter", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy 5 times (scale reference for HBM byte counters)")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-rates", type=str, default="160,20",
                    help="Poisson request rates (req/s), run one after the other on the same sessions; the "
                         "default pair is near the measured capacity (~175 req/s on the 256 sessions) and light "
                         "load")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch

    from collisiondetection_amd import _native as N
    from collisiondetection_amd import api, scenes

    dev = torch.device("cuda", 0)
    scene = scenes.cloth(args.grid) if args.workload == "cloth" else scenes.crumpled_sheet(args.grid)
    F, V = len(scene["faces"]), len(scene["qstart"])
    log(f"workload: {args.workload} {args.grid}^2, {V} vertices, {F} triangles, eta {scene['eta']}")
    ctx = api.Context(0)
    ctx.bind_stream()
    faces = api.to_device(scene["faces"], torch.int32, ctx).reshape(-1)
    qs = api.to_device(scene["qstart"], torch.float64, ctx).reshape(-1)
    qe = api.to_device(scene["qend"], torch.float64, ctx).reshape(-1)
    eta = scene["eta"]
    bufs = {}

    def alloc(cap):
        bufs.update(cap=cap, vk=torch.empty(4 * cap, dtype=torch.int32, device=dev),
                    vt=torch.empty(cap, dtype=torch.float64, device=dev),
                    ek=torch.empty(4 * cap, dtype=torch.int32, device=dev),
                    et=torch.empty(cap, dtype=torch.float64, device=dev))

    alloc(4 * F)
    cnt = [ctypes.c_int64(0) for _ in range(4)]
    L = N.lib()

    def step():
        b = bufs
        rc = L.ccd_detect(ctx.ptr, faces.data_ptr(), F, qs.data_ptr(), qe.data_ptr(), V, eta, eta, None,
                          b["vk"].data_ptr(), b["vt"].data_ptr(), b["cap"], ctypes.byref(cnt[0]), b["ek"].data_ptr(),
                          b["et"].data_ptr(), b["cap"], ctypes.byref(cnt[1]), ctypes.byref(cnt[2]),
                          ctypes.byref(cnt[3]))
        if rc == N.CCD_E_CAPACITY:   # collision buffers too small (warmup only): grow and repeat
            alloc(max(cnt[0].value, cnt[1].value) * 5 // 4 + 1024)
            return step()
        ctx.check(rc)

    def run_detect():
        """Warmup, work accounting and the timed region; returns the figures."""
        for i in range(args.warmup):
            step()
            log(f"warmup {i + 1}/{args.warmup}: {cnt[2].value} VF + {cnt[3].value} EE "
                f"candidates, {cnt[0].value} + {cnt[1].value} collisions")
        torch.cuda.synchronize(dev)

        # Work accounting for the narrowphase roofline (outside the timed region).
        pv, pe = ctypes.c_void_p(), ctypes.c_void_p()
        nv, ne = ctypes.c_int64(), ctypes.c_int64()
        ctx.check(L.ccd_last_candidates(ctx.ptr, ctypes.byref(pv), ctypes.byref(nv), ctypes.byref(pe),
                                        ctypes.byref(ne)))
        wv = (ctypes.c_int64 * N.NSTAGES)()
        we = (ctypes.c_int64 * N.NSTAGES)()
        ctx.check(L.ccd_narrowphase_work(ctx.ptr, qs.data_ptr(), qe.data_ptr(), V, pv, nv.value, pe, ne.value, eta,
                                         wv, we))
        fl = (ctypes.c_int64 * N.NSTAGES)()
        L.ccd_stage_flops(fl)
        flops_vf = sum(wv[i] * fl[i] for i in range(N.NSTAGES))
        flops_ee = sum(we[i] * fl[i] for i in range(N.NSTAGES))

        # Timed region: K steps, HIP-event stage timers on the context's stream.
        L.ccd_set_profiling(ctx.ptr, 1)
        L.ccd_stage_times(ctx.ptr, None, None, 1)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize(dev)
        wall = time.perf_counter() - t0
        ms = (ctypes.c_double * N.NTIMERS)()
        calls = (ctypes.c_int64 * N.NTIMERS)()
        L.ccd_stage_times(ctx.ptr, ms, calls, 1)
        L.ccd_set_profiling(ctx.ptr, 0)
        stage_ms = {N.TIMER_NAMES[i]: ms[i] / max(calls[i], 1) for i in range(N.NTIMERS)}
        log("stage ms/step:", json.dumps({k: round(v, 4) for k, v in stage_ms.items()}))
        return {"wall": wall, "stage_ms": stage_ms, "nv": nv.value, "ne": ne.value, "flops_vf": flops_vf,
                "flops_ee": flops_ee, "hits": (cnt[0].value, cnt[1].value)}

    res = run_detect()
    wall, stage_ms, flops_vf, flops_ee = res["wall"], res["stage_ms"], res["flops_vf"], res["flops_ee"]
    nv, ne = ctypes.c_int64(res["nv"]), ctypes.c_int64(res["ne"])
    ms_per_step = 1000.0 * wall / args.steps
    value = F / (wall / args.steps)
    nstencils = nv.value + ne.value

    out_workload = (f"cloth self-CCD {args.grid}x{args.grid} ({V} V, {F} tris), eta 1e-6, seed 1"
                    if args.workload == "cloth" else
                    f"crumpled sheet self-CCD {args.grid}x{args.grid} ({V} V, {F} tris), eta 1e-7, seed 3")
    # Roofline of the dominant kernel.
    dom = max((k for k in stage_ms if k != "pipeline"), key=stage_ms.get)
    if dom in ("narrow_vf", "narrow_ee"):
        flops = flops_vf if dom == "narrow_vf" else flops_ee
        achieved = flops / (stage_ms[dom] * 1e-3) / 1e12
        roof = {"kernel": "k_narrow2_queue<%s>" % ("VF" if dom == "narrow_vf" else "EE"), "bound": "fp64",
                "pipe": "FP64 VALU (no dense contraction, so no MFMA; peak is MI355X's FP64 figure)",
                "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": None, "algorithmic_flops": flops}
    else:
        # Broadphase stages are bandwidth/latency bound: algorithmic bytes per swept triangle =
        # 12 (face) + 48 V/F (start + end positions) + 16 S/F (unique stencil writes).
        algo = 12 * F + 48 * V + 16 * nstencils
        achieved = algo / (stage_ms[dom] * 1e-3) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": algo}
    # HBM traffic of the dominant kernel per launch, from the committed rocprofv3 counter passes on this
    # workload (FETCH_SIZE x 2 + WRITE_SIZE, the factors a plain device copy gives); null elsewhere.
    try:
        prof = json.load(open(os.path.join(ROOT, "profiles", "cloth256_traffic.json")))
        key = roof["kernel"].replace("k_narrow2_queue<VF>", "k_narrow2_queue<true, 256, ccd::NoCount>").replace(
            "k_narrow2_queue<EE>", "k_narrow2_queue<false, 256, ccd::NoCount>")
        if prof["workload"] == out_workload and key in prof["kernels"]:
            kp = prof["kernels"][key]
            roof["traffic"] = kp["hbm_bytes"]
            roof["traffic_source"] = "profiles/cloth256_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"
            if "fp64_wave_instructions" in kp and roof.get("unit") == "TFLOP/s":
                # Executed FP64 wavefront instructions (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 of the same
                # profile) against the issue peak: 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction.
                issue_peak = 1024 * 2.4e9 / 4
                roof["fp64_issue_frac"] = kp["fp64_wave_instructions"] / (kp["avg_ms"] * 1e-3) / issue_peak
    except (OSError, ValueError, KeyError):
        pass
    roof["stage_ms"] = {k: round(v, 5) for k, v in stage_ms.items()}
    roof["narrowphase_tflops"] = {
        "vf": flops_vf / max(stage_ms["narrow_vf"], 1e-9) / 1e9, "ee": flops_ee / max(stage_ms["narrow_ee"], 1e-9) / 1e9}

    out = {
        "metric": "swept tris/sec on 1 GPU, batched vs plain, exact sets; p50/p99 CCD query latency",
        "value": value, "unit": "swept tris/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": out_workload,
                   "global_batch": 1, "parallelism": "single GPU"},
        "pair_tests_per_s": nstencils / ((stage_ms["narrow_vf"] + stage_ms["narrow_ee"]) * 1e-3),
        "candidates": {"vf": nv.value, "ee": ne.value}, "collisions": {"vf": res["hits"][0], "ee": res["hits"][1]},
        "roots": "isolated in [0,1] (derivative bracketing + Newton), in place of the reference's rpoly",
        "roofline": roof,
    }
    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize(dev)
        out["copy_probe_bytes"] = {"read": 256 << 20, "written": 256 << 20, "dispatches": 5}
    if not args.no_separating_plane:
        out["separating_plane"] = bench_separating_plane(ctx, scene, qs, qe, max(3, args.steps // 2), log)
    if not args.no_velocity_filter:
        out["velocity_filter"] = bench_velocity_filter(ctx, log)
    if not args.no_batched:
        out["batched"] = bench_batched(max(3, args.steps // 2), log)
    if not args.no_served:
        out["served"] = bench_served(args.serve_seconds, [float(r) for r in args.serve_rates.split(",")], log)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread) ...")
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
