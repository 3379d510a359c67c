"""Benchmark: swept triangles per second of the swept-primitive CCD path on one MI355X.

Headline workload: the batched one -- 512 independent two-mesh instances (icospheres or noisy grids,
log-uniform 80-20,480 triangles per mesh, 4,018,680 triangles in all), each mesh moved by a random
rigid motion plus N(0, 0.005^2) per vertex, seed 2, eta = 1e-6, FP64.  One step = the whole batch
through the reference's KDOPBroadPhase (leaf bounds, tree, traversal, stencil set) and CTCDNarrowPhase
on every candidate stencil, as one pipeline replayed from a captured HIP graph, positions already
resident in HBM; the colliding stencils and their times of impact come back to the host.  Beside it,
"plain": the same instances one ccd_detect call each (the reference runs one mesh per call).

Printed: one JSON line with the metric, the batched and plain rates, the roofline of the pipeline's
dominant kernel (HIP-event stage timers on the context's stream), host description and CPU baselines
(the oracle's reference-structured path on a stratified sample of the instances, 1 thread and all
cores), plus secondary legs, each switched by --legs: the 256^2 draped cloth (cloth), the 1024^2
crumpled sheet (sheet), the separating-plane narrowphase on the cloth's candidates (separating_plane),
the velocity filter (velocity_filter) and the serving loop over 256 resident sessions (served).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# FP64 wavefront-instruction issue peak: 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 FP64 instruction.
FP64_ISSUE_PEAK = 1024 * 2.4e9 / 4
ETA_BATCH = 1e-6
ALL_LEGS = ("plain", "cloth", "sheet", "separating_plane", "velocity_filter", "served")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0))
    threads = min(usable, int(os.environ.get("OMP_NUM_THREADS", usable)))
    return {"cpu_model": model, "usable_cpus": usable, "logical_cpus": os.cpu_count(), "baseline_threads": threads}


def load_profile(name, workload):
    """Per-kernel counter figures from a committed rocprofv3 summary (profiles/<name>.json), if it was
    made on this same workload."""
    try:
        prof = json.load(open(os.path.join(ROOT, "profiles", name)))
    except (OSError, ValueError):
        return None
    return prof if prof.get("workload") == workload else None


def roofline(kernel, stage_ms, flops, prof, launches_per_step=1):
    """FP64 roofline of one narrowphase kernel: algorithmic flops (the survey's coefficient-construction
    count of the sub-tests the reference executes) over its HIP-event time; plus, from the committed
    counter profile of the same workload, the FP64 work the kernel issues (SQ_INSTS_VALU_*_F64 wavefront
    instructions x 64 lanes, FMA counted twice: root finding included, lanes masked off by divergence
    counted too) and the HBM traffic per launch."""
    t = stage_ms * 1e-3 / launches_per_step
    achieved = flops / launches_per_step / t / 1e12
    roof = {"kernel": kernel, "bound": "fp64", "pipe": "FP64 VALU (no dense contraction, so no MFMA)",
            "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None, "algorithmic_flops_per_launch": flops / launches_per_step, "launch_ms": t * 1e3}
    kp = (prof or {}).get("kernels", {}).get(kernel)
    if kp:
        if "hbm_bytes" in kp:
            roof["traffic"] = kp["hbm_bytes"]
            roof["traffic_source"] = prof.get("source")
        if "fp64_wave_instructions" in kp:
            ex = kp["fp64_issue_flops"] / (kp["avg_ms"] * 1e-3) / 1e12
            roof["executed"] = {"achieved": ex, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": ex / FP64_PEAK_TFLOPS,
                                "issue_frac": kp["fp64_wave_instructions"] / (kp["avg_ms"] * 1e-3) / FP64_ISSUE_PEAK,
                                "source": prof.get("source")}
    return roof


def work_flops(ctx, qs_ptr, qe_ptr, V, eta):
    """Algorithmic flops of the last detection's narrowphase on this context (VF, EE)."""
    from collisiondetection_amd import _native as N

    L = N.lib()
    pv, pe = ctypes.c_void_p(), ctypes.c_void_p()
    nv, ne = ctypes.c_int64(), ctypes.c_int64()
    ctx.check(L.ccd_last_candidates(ctx.ptr, ctypes.byref(pv), ctypes.byref(nv), ctypes.byref(pe), ctypes.byref(ne)))
    wv = (ctypes.c_int64 * N.NSTAGES)()
    we = (ctypes.c_int64 * N.NSTAGES)()
    ctx.check(L.ccd_narrowphase_work(ctx.ptr, qs_ptr, qe_ptr, V, pv, nv.value, pe, ne.value, eta, wv, we))
    fl = (ctypes.c_int64 * N.NSTAGES)()
    L.ccd_stage_flops(fl)
    return (sum(wv[i] * fl[i] for i in range(N.NSTAGES)), sum(we[i] * fl[i] for i in range(N.NSTAGES)),
            nv.value, ne.value)


def stage_profile(ctx, run, steps):
    """Mean per-stage milliseconds (HIP events on the context's stream) over `steps` calls of run()."""
    import torch

    from collisiondetection_amd import _native as N

    L = N.lib()
    L.ccd_set_profiling(ctx.ptr, 1)
    L.ccd_stage_times(ctx.ptr, None, None, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    ms = (ctypes.c_double * N.NTIMERS)()
    calls = (ctypes.c_int64 * N.NTIMERS)()
    L.ccd_stage_times(ctx.ptr, ms, calls, 1)
    L.ccd_set_profiling(ctx.ptr, 0)
    return {N.TIMER_NAMES[i]: ms[i] / max(calls[i], 1) for i in range(N.NTIMERS)}, wall


def cpu_baseline_batched(inst, threads, seed=0, strata=8, per_stratum=2, stencil_sample=4000):
    """The oracle's reference-structured CPU path (top-down k-DOP tree, recursive traversal into
    std::set, then the per-stencil narrowphase with the reference's own root finder) on a stratified
    sample of the instances: instances sorted by size into `strata` groups, `per_stratum` drawn from
    each; per drawn instance the broadphase in full and the narrowphase on a random sample of its
    stencils, extrapolated per stencil; the group's time = its size / drawn x the drawn instances'
    times.  threads > 1: the same restatement with the traversal's subtree pairs and the stencils spread
    over OpenMP threads."""
    import numpy as np

    import oracle

    rf = oracle.use_default_rootfinder()
    oracle.set_threads(threads)
    rng = np.random.default_rng(seed)
    sizes = np.array([len(s["faces"]) for s in inst])
    groups = np.array_split(np.argsort(sizes, kind="stable"), strata)
    total, spent, drawn = 0.0, 0.0, []
    try:
        for g in groups:
            pick = rng.choice(g, size=min(per_stratum, len(g)), replace=False)
            t_g = 0.0
            for k in pick:
                s = inst[int(k)]
                h = oracle.Hist.two_entry(s["qstart"], s["qend"])
                t0 = time.perf_counter()
                vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree_set" if threads == 1 else "tree_parallel")
                tb = time.perf_counter() - t0
                tn = 0.0
                for keys, is_ee in ((vf, False), (ee, True)):
                    if not len(keys):
                        continue
                    m = min(len(keys), stencil_sample)
                    sub = keys[rng.choice(len(keys), size=m, replace=False)]
                    t0 = time.perf_counter()
                    oracle.narrowphase(h, sub, s["eta"], is_ee)
                    tn += (time.perf_counter() - t0) * len(keys) / m
                t_g += tb + tn
                spent += tb
                drawn.append(int(k))
            total += t_g * len(g) / len(pick)
    finally:
        oracle.set_threads(1)
    F = int(sizes.sum())
    return {"value": F / total, "unit": "swept tris/s", "cores": threads, "kind": "port",
            "sample": (f"oracle (root finder '{rf}', {threads} thread{'s' if threads > 1 else ''}) on {len(drawn)} of "
                       f"{len(inst)} instances, {per_stratum} drawn from each of {strata} size strata; broadphase in "
                       f"full, narrowphase on <= {stencil_sample} random stencils per kind and instance, extrapolated: "
                       f"{total:.1f} s per batch")}


def cpu_baseline_scene(scene, threads=1, sample_stencils=60000, seed=0):
    """The oracle's reference-structured CPU path on one mesh: the broadphase in full and the
    narrowphase on a random sample of its stencils, extrapolated per stencil."""
    import numpy as np

    import oracle

    rf = oracle.use_default_rootfinder()
    oracle.set_threads(threads)
    try:
        h = oracle.Hist.two_entry(scene["qstart"], scene["qend"])
        t0 = time.perf_counter()
        vf, ee = oracle.candidates(scene["faces"], h, scene["eta"], method="tree_set" if threads == 1 else "tree_parallel")
        t_broad = time.perf_counter() - t0
        rng = np.random.default_rng(seed)
        nv = min(len(vf), sample_stencils * len(vf) // max(1, len(vf) + len(ee)))
        ne = min(len(ee), sample_stencils - nv)
        t_narrow = 0.0
        for keys, m, is_ee in ((vf, nv, False), (ee, ne, True)):
            if m:
                sub = keys[rng.choice(len(keys), size=m, replace=False)]
                t0 = time.perf_counter()
                oracle.narrowphase(h, sub, scene["eta"], is_ee)
                t_narrow += (time.perf_counter() - t0) * len(keys) / m
    finally:
        oracle.set_threads(1)
    total = t_broad + t_narrow
    return {"value": len(scene["faces"]) / total, "unit": "swept tris/s", "cores": threads, "kind": "port",
            "sample": (f"oracle (root finder '{rf}', {threads} thread{'s' if threads > 1 else ''}); full broadphase "
                       f"({t_broad:.2f} s, {len(vf)} VF + {len(ee)} EE stencils) + narrowphase on {nv} VF and {ne} EE "
                       f"sampled stencils, extrapolated to all: {total:.1f} s per step")}


def bench_detect(scene, name, steps, warmup, log):
    """One mesh through ccd_detect (broadphase + narrowphase): wall time per step and HIP-event stages."""
    import torch

    from collisiondetection_amd import _native as N
    from collisiondetection_amd import api

    L = N.lib()
    ctx = api.Context(0)
    ctx.bind_stream()
    dev = ctx.torch_device
    F, V = len(scene["faces"]), len(scene["qstart"])
    faces = api.to_device(scene["faces"], torch.int32, ctx).reshape(-1)
    qs = api.to_device(scene["qstart"], torch.float64, ctx).reshape(-1)
    qe = api.to_device(scene["qend"], torch.float64, ctx).reshape(-1)
    eta = scene["eta"]
    bufs = {}
    cnt = [ctypes.c_int64(0) for _ in range(4)]

    def alloc(cap):
        bufs.update(cap=cap, vk=torch.empty(4 * cap, dtype=torch.int32, device=dev),
                    vt=torch.empty(cap, dtype=torch.float64, device=dev),
                    ek=torch.empty(4 * cap, dtype=torch.int32, device=dev),
                    et=torch.empty(cap, dtype=torch.float64, device=dev))

    alloc(max(1024, F))

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

    for i in range(warmup):
        step()
        log(f"{name}: warmup {i + 1}/{warmup}: {cnt[2].value} VF + {cnt[3].value} EE candidates, "
            f"{cnt[0].value} + {cnt[1].value} collisions")
    torch.cuda.synchronize(dev)
    fv, fe, nv, ne = work_flops(ctx, qs.data_ptr(), qe.data_ptr(), V, eta)
    stage_ms, wall = stage_profile(ctx, step, steps)
    log(f"{name}: {wall * 1e3:.2f} ms/step; stages", json.dumps({k: round(v, 3) for k, v in stage_ms.items()}))
    out = {"triangles": F, "vertices": V, "ms_per_step": wall * 1e3, "swept_tris_per_s": F / wall,
           "pair_tests_per_s": (nv + ne) / ((stage_ms["narrow_vf"] + stage_ms["narrow_ee"]) * 1e-3),
           "candidates": {"vf": nv, "ee": ne}, "collisions": {"vf": cnt[0].value, "ee": cnt[1].value},
           "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
           "narrowphase_tflops": {"vf": fv / max(stage_ms["narrow_vf"], 1e-9) / 1e9,
                                  "ee": fe / max(stage_ms["narrow_ee"], 1e-9) / 1e9}}
    return out, (ctx, qs, qe, stage_ms, fv, fe)


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


def bench_served(seconds, rates, log, sessions=256, budget_gib=32, max_inflight=16):
    """The served workload: `sessions` device-resident sessions (log-uniform 1k-500k triangles, seed 4),
    driven by the C++ load generator (collisiondetection_amd/serve_loadgen): for each rate, Poisson
    arrivals at that many requests/s for `seconds`, each to a uniform session, supplying new end
    positions = previous end + a fresh N(0, 0.002^2) draw per coordinate; latency from the scheduled
    arrival to the result on the host.  A run stops offering once two seconds' worth of requests wait
    (overload), then drains."""
    import subprocess
    import tempfile

    import numpy as np

    from collisiondetection_amd import scenes

    exe = os.path.join(ROOT, "collisiondetection_amd", "serve_loadgen")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build with collisiondetection_amd._build.build()")
    sess = scenes.served_sessions(sessions)
    nf = np.array([len(s["faces"]) for s in sess])
    log(f"served workload: {len(sess)} sessions, {int(nf.sum())} resident triangles")
    runs = []
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "sessions.bin")
        with open(path, "wb") as fh:
            np.array([len(sess)], np.int64).tofile(fh)
            for s in sess:
                np.array([len(s["faces"]), len(s["qstart"])], np.int64).tofile(fh)
                np.array([s["eta"]], np.float64).tofile(fh)
                np.ascontiguousarray(s["faces"], np.int32).tofile(fh)
                np.ascontiguousarray(s["qstart"], np.float64).tofile(fh)
        for rate in rates:
            cmd = [exe, path, str(rate), str(seconds), "--inflight", str(max_inflight), "--budget-gib",
                   str(budget_gib), "--threads", "8", "--seed", "4"]
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=seconds * 4 + 300)
            if r.returncode != 0:
                raise RuntimeError(f"load generator failed with {r.returncode}")
            out = json.loads(r.stdout.strip().splitlines()[-1])
            log("served:", json.dumps(out))
            runs.append(out)
    return {"sessions": len(sess), "resident_triangles": int(nf.sum()), "budget_gib": budget_gib,
            "max_inflight": max_inflight, "runs": runs}


def bench_plain(b, inst, steps, log):
    """The batched workload's instances one ccd_detect call each (per-instance launches and host
    synchronisation, as the reference runs one mesh per call): wall time per pass over all instances."""
    import torch

    from collisiondetection_amd import _native as N
    from collisiondetection_amd import api

    ctx, L = b.ctx, N.lib()
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
    for _ in range(steps):
        hits = plain()
    torch.cuda.synchronize(dev)
    wall = (time.perf_counter() - t0) / steps
    log(f"plain: {wall * 1e3:.1f} ms per pass over {len(inst)} instances, {hits} collisions")
    return wall, hits


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--legs", type=str, default=",".join(ALL_LEGS),
                    help="secondary legs to run, comma separated (%s), or 'none'" % ", ".join(ALL_LEGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune-cache", type=str, default=os.path.join(ROOT, "collisiondetection_amd", "queue_shapes.json"),
                    help="JSON cache of tuned narrowphase shapes per GPU and workload shape")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy 5 times (scale reference for HBM byte counters)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rates", type=str, default="2000,250",
                    help="Poisson request rates (req/s), each on freshly opened sessions: the configured 2,000 "
                         "(above one GPU's capacity for these sessions: the run stops at overload) and a "
                         "sustainable one")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    legs = set() if args.legs == "none" else {x.strip() for x in args.legs.split(",") if x.strip()}
    unknown = legs - set(ALL_LEGS)
    if unknown:
        print(f"unknown legs: {sorted(unknown)}", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from collisiondetection_amd import api, scenes

    dev = torch.device("cuda", 0)
    host = host_info()
    log("host:", json.dumps(host))
    inst = scenes.batched_instances()
    b = api.Batch(inst)
    ctx = b.ctx
    F, V = b.F, b.V
    workload = (f"batched: {b.ninst} two-mesh instances (icospheres / noisy grids, log-uniform 80-20,480 tris per "
                f"mesh), {F} tris, rigid motion + N(0, 0.005^2), seed 2, eta 1e-6")
    log(f"workload: {workload}")
    # Queue shape and stencil grouping per kind from the committed tuning cache for this workload shape
    # (tuned now, and stored, when the cache has no entry for this GPU and workload).
    shape, how = b.tune(ETA_BATCH, ETA_BATCH, cache=args.tune_cache)
    log(f"narrowphase shape ({how}): {shape}")
    for i in range(args.warmup):   # the first call runs synchronously and captures the pipeline
        r = b.detect(ETA_BATCH, ETA_BATCH, graph=True)
        log(f"warmup {i + 1}/{args.warmup}: {r['vf_candidates']} VF + {r['ee_candidates']} EE candidates, "
            f"{len(r['vf'])} + {len(r['ee'])} collisions")
    torch.cuda.synchronize(dev)
    # Timed region: K graph replays of the whole batch, results back on the host.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = b.detect(ETA_BATCH, ETA_BATCH, graph=True)
    torch.cuda.synchronize(dev)
    wall = (time.perf_counter() - t0) / args.steps
    value = F / wall
    hits = len(r["vf"]) + len(r["ee"])
    log(f"batched (graph replay): {wall * 1e3:.2f} ms/step, {value / 1e6:.3f}M swept tris/s")
    # The same pipeline launched stage by stage (HIP-event stage timers on the context's stream), and the
    # algorithmic work of its candidates.
    fv, fe, nv, ne = work_flops(ctx, b.qs.data_ptr(), b.qe.data_ptr(), V, ETA_BATCH)
    stage_ms, wall_sync = stage_profile(ctx, lambda: b.detect(ETA_BATCH, ETA_BATCH, graph=False), max(2, args.steps // 3))
    log(f"batched (stage-timed synchronous run): {wall_sync * 1e3:.2f} ms/step; stages",
        json.dumps({k: round(v, 3) for k, v in stage_ms.items()}))
    prof = load_profile("batched512_counters.json", workload)
    dom = max(("narrow_vf", "narrow_ee"), key=stage_ms.get)
    # The edge-edge stage is two kernels (first pass, then the deferred stencils): counted together.
    kern = "k_narrow2_queue<VF>" if dom == "narrow_vf" else "k_narrow2_queue_ee_shared+k_ee_deferred"
    roof = roofline(kern, stage_ms[dom], fv if dom == "narrow_vf" else fe, prof)
    roof["stage_ms"] = {k: round(v, 5) for k, v in stage_ms.items()}
    roof["narrowphase_tflops"] = {"vf": fv / max(stage_ms["narrow_vf"], 1e-9) / 1e9,
                                  "ee": fe / max(stage_ms["narrow_ee"], 1e-9) / 1e9}
    out = {
        "metric": "swept tris/sec on 1 GPU, batched vs plain, exact sets; p50/p99 CCD query latency",
        "value": value, "unit": "swept tris/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "instances": b.ninst, "triangles": F, "vertices": V,
                   "global_batch": b.ninst, "parallelism": "single GPU"},
        "pair_tests_per_s": (nv + ne) / ((stage_ms["narrow_vf"] + stage_ms["narrow_ee"]) * 1e-3),
        "candidates": {"vf": nv, "ee": ne}, "collisions": {"vf": len(r["vf"]), "ee": len(r["ee"])},
        "roots": "isolated in [0,1] (derivative bracketing + Newton), in place of the reference's rpoly",
        "stage_timed_ms_per_step": wall_sync * 1e3,
        "narrowphase_shape": {"source": how, "vf": shape[0:3] + [shape[6]], "ee": shape[3:6] + [shape[7]],
                              "fields": ["block", "tile", "waves_per_cu", "grouping"]},
        "roofline": roof,
        "host": host,
    }
    if "plain" in legs:
        t_plain, h_plain = bench_plain(b, inst, max(1, args.steps // 4), log)
        out["plain"] = {"ms_per_step": t_plain * 1e3, "swept_tris_per_s": F / t_plain,
                        "batched_vs_plain": t_plain / wall, "collisions_equal": h_plain == hits}
    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize(dev)
        out["copy_probe_bytes"] = {"read": 256 << 20, "written": 256 << 20, "dispatches": 5}
    cloth = None
    if "cloth" in legs or "separating_plane" in legs:
        cloth = scenes.cloth(256)
    if "cloth" in legs:
        leg, _ = bench_detect(cloth, "cloth 256^2", args.steps, args.warmup, log)
        leg["workload"] = "cloth self-CCD 256x256 (65536 V, 130050 tris), eta 1e-6, seed 1"
        out["cloth"] = leg
    if "sheet" in legs:
        sheet = scenes.crumpled_sheet(1024)
        leg, _ = bench_detect(sheet, "sheet 1024^2", 1, 1, log)
        leg["workload"] = "crumpled sheet self-CCD 1024x1024 (1048576 V, 2093058 tris), eta 1e-7, seed 3"
        out["sheet"] = leg
        del sheet
        torch.cuda.empty_cache()
    if "separating_plane" in legs:
        qs = api.to_device(cloth["qstart"], torch.float64, ctx).reshape(-1)
        qe = api.to_device(cloth["qend"], torch.float64, ctx).reshape(-1)
        out["separating_plane"] = bench_separating_plane(ctx, cloth, qs, qe, max(3, args.steps // 2), log)
    if "velocity_filter" in legs:
        out["velocity_filter"] = bench_velocity_filter(ctx, log)
    if "served" in legs:
        out["served"] = bench_served(args.serve_seconds, [float(x) for x in args.serve_rates.split(",")], log)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread, stratified sample of the instances) ...")
        out["cpu_baseline"] = cpu_baseline_batched(inst, 1)
        log("cpu baseline:", json.dumps(out["cpu_baseline"]))
        if host["baseline_threads"] > 1:
            log(f"cpu baseline (oracle, {host['baseline_threads']} threads) ...")
            out["cpu_baseline_all_cores"] = cpu_baseline_batched(inst, host["baseline_threads"])
            log("cpu baseline, all cores:", json.dumps(out["cpu_baseline_all_cores"]))
        if "cloth" in legs:
            out["cloth"]["cpu_baseline"] = cpu_baseline_scene(cloth, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
