"""The product's FP64 core (collisiondetection_amd/csrc/ccd_core.h), compiled for the host, against the
oracle: the algorithm itself, before any device arithmetic is involved."""
import ctypes

import numpy as np
import pytest

import oracle
from collisiondetection_amd import scenes
from tests.helpers import (compare_results, core_find_intervals, core_narrowphase, core_primitive, core_roots01, dp,
                           exact_arbiters)
from tests.test_oracle import ETA, TESTCTCD, crossing_cases, expected_testctcd

KIND = {"vf": 0, "ee": 1, "ve": 2, "vv": 3}


def test_core_testctcd_cases(core_harness):
    exp = expected_testctcd()
    for k in ("vv", "ve", "vf", "ee"):
        r, t = core_primitive(core_harness, KIND[k], TESTCTCD[k], ETA)
        o = {"vv": oracle.vertex_vertex_ctcd, "ve": oracle.vertex_edge_ctcd, "vf": oracle.vertex_face_ctcd,
             "ee": oracle.edge_edge_ctcd}[k](TESTCTCD[k], ETA)
        assert r and o[0]
        assert abs(t - o[1]) < 1e-12
        assert abs(t - exp[k]) < 1e-10
    # closed-form (degree <= 2) answers are bit-identical
    assert core_primitive(core_harness, 3, TESTCTCD["vv"], ETA)[1] == exp["vv"]
    assert core_primitive(core_harness, 2, TESTCTCD["ve"], ETA)[1] == exp["ve"]


def test_core_crossing_cases(core_harness):
    for kind, q, hit, t in crossing_cases(ETA):
        r, tt = core_primitive(core_harness, KIND[kind], q, ETA)
        assert r == hit, kind
        if hit:
            assert abs(tt - t) < 1e-10


def test_core_quadratic_intervals_bit_exact(core_harness):
    rng = np.random.default_rng(3)
    for _ in range(3000):
        op = rng.normal(size=3)
        op[rng.random(3) < 0.15] = 0.0
        for pos in (True, False):
            a = core_find_intervals(core_harness, op, pos)
            b, _, _ = oracle.find_intervals(op, pos)
            assert a == b, (op, pos, a, b)


@pytest.mark.parametrize("deg", [3, 4, 6])
def test_core_intervals_vs_oracle(core_harness, deg):
    rng = np.random.default_rng(deg)
    bad = 0
    for _ in range(3000):
        op = rng.normal(size=deg + 1) * np.exp(2 * rng.normal(size=deg + 1))
        for pos in (True, False):
            a = core_find_intervals(core_harness, op, pos)
            b, _, m = oracle.find_intervals(op, pos)
            # union of the intervals: equal up to the root-finder accuracy
            ua = _union(a)
            ub = _union(b)
            if len(ua) != len(ub) or any(abs(x - y) > 1e-9 for p, q in zip(ua, ub) for x, y in zip(p, q)):
                if m >= 1e-12:
                    bad += 1
    assert bad == 0


def _union(iv):
    iv = sorted(iv)
    out = []
    for l, u in iv:
        if out and l <= out[-1][1]:
            out[-1] = (out[-1][0], max(out[-1][1], u))
        else:
            out.append((l, u))
    return out


def test_core_roots_vs_mpmath(core_harness):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    rng = np.random.default_rng(21)
    for _ in range(300):
        n = int(rng.integers(3, 7))
        # polynomials with known roots, some inside [0,1]
        roots = rng.uniform(-0.5, 1.5, size=n)
        c = np.poly(roots) * rng.uniform(0.5, 2)
        got = core_roots01(core_harness, c)
        ref = sorted(float(mp.re(z)) for z in mp.polyroots([mp.mpf(float(x)) for x in c], maxsteps=300, extraprec=300)
                     if abs(mp.im(z)) < 1e-20 and 0 <= mp.re(z) <= 1)
        gaps = np.diff(np.sort(roots)) if n > 1 else [1]
        if min(gaps) < 1e-6:
            continue
        assert len(got) == len(ref)
        # forward error of a root of the double polynomial: eps * sum|c_i||r|^(n-i) / |p'(r)|
        for g, r in zip(got, ref):
            cond = sum(abs(ci) * abs(r) ** (n - i) for i, ci in enumerate(c)) / abs(np.polyval(np.polyder(c), r))
            assert abs(g - r) <= 8 * 2.2e-16 * cond + 1e-15, (g, r, cond)


@pytest.mark.parametrize("scene", ["two_spheres", "converging", "cloth64"])
def test_core_narrowphase_vs_oracle(core_harness, scene):
    s = {"two_spheres": scenes.two_spheres, "converging": scenes.converging_spheres,
         "cloth64": lambda: scenes.cloth(64, drop=0.05, noise=0.01)}[scene]()
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
    for keys, is_ee in ((vf, False), (ee, True)):
        rec = oracle.narrowphase(h, keys, s["eta"], is_ee)
        hit, t, sub = core_narrowphase(core_harness, s["qstart"], s["qend"], keys, is_ee, s["eta"])
        rep = compare_results(scene, hit, t, rec, arbiter=exact_arbiters(h, keys, s["eta"], is_ee))
        print(rep)
        assert not rep["hard"], rep
        assert not rep["t_bad_hard"], rep
        both = hit & rec["hit"].astype(bool)
        assert np.array_equal(sub[both], rec["subtest"][both])


# ------------------------------------------------------------------------------------------------
# Reference-root mode (ccd_rpoly.h): the reference's Jenkins-Traub solver restated.
# ------------------------------------------------------------------------------------------------
def _core_fn(H, name, restype=ctypes.c_double):
    f = getattr(H, name)
    f.restype = restype
    f.argtypes = [ctypes.c_double]
    return f


def test_core_log_exp_correctly_rounded(core_harness):
    """The restated solver's log/exp are correctly rounded (checked against 200-bit mpmath), so host and
    device agree bit for bit; the C library's log/exp agree with them except where it misrounds."""
    import math

    import mpmath

    mpmath.mp.prec = 200
    lg, ex = _core_fn(core_harness, "core_log_cr"), _core_fn(core_harness, "core_exp_cr")
    rng = np.random.default_rng(11)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 1500)), rng.uniform(0.5, 2.0, 1500),
                         2.0 ** rng.integers(-1070, 1020, 100).astype(float), [4.9e-324, 1e-310, 1.0, 2.0]])
    for x in xs:
        assert lg(x) == float(mpmath.log(mpmath.mpf(x))), x
    ys = np.concatenate([rng.uniform(-700, 700, 1500), rng.uniform(-1, 1, 1500), [0.0, 1.0, -1.0, 700.0]])
    for y in ys:
        assert ex(y) == float(mpmath.exp(mpmath.mpf(y))), y
    # glibc agreement rate (informational): misrounded cases are rare
    agree = np.mean([lg(x) == math.log(x) for x in xs])
    assert agree > 0.99


def test_core_log_exp_fast_paths(core_harness):
    """The fast paths of log/exp (series tail in double, rounding settled by an error bound) stay well
    inside the bound their rounding test assumes (measured against 200-bit mpmath), and with the
    fallback the results equal the full double-double evaluation on every input."""
    import mpmath

    mpmath.mp.prec = 200
    H = core_harness
    for n in ("core_log_cr", "core_exp_cr", "core_log_cr_full", "core_exp_cr_full"):
        _core_fn(H, n)
    H.core_fast_eps.restype = ctypes.c_double
    eps = H.core_fast_eps()
    H.core_log_fast_dd.argtypes = [ctypes.c_double, dp]
    H.core_exp_fast_dd.argtypes = [ctypes.c_double, dp]
    H.core_exp_fast_dd.restype = ctypes.c_int
    out = np.zeros(2)
    rng = np.random.default_rng(12)
    worst = 0.0
    xs = np.concatenate([np.exp(rng.uniform(-690, 700, 3000)), rng.uniform(0.5, 2.0, 3000),
                         1 + rng.uniform(-1e-3, 1e-3, 1000), 1 + rng.uniform(-1e-12, 1e-12, 1000)])
    for x in xs:
        H.core_log_fast_dd(x, out.ctypes.data_as(dp))
        v = mpmath.log(mpmath.mpf(x))
        if v != 0:
            worst = max(worst, float(abs((mpmath.mpf(out[0]) + mpmath.mpf(out[1]) - v) / v)))
    ys = np.concatenate([rng.uniform(-700, 700, 3000), rng.uniform(-1, 1, 3000), rng.uniform(-1e-6, 1e-6, 1000)])
    for y in ys:
        k = H.core_exp_fast_dd(y, out.ctypes.data_as(dp))
        v = mpmath.exp(mpmath.mpf(y)) * mpmath.mpf(2) ** (-k)
        worst = max(worst, float(abs((mpmath.mpf(out[0]) + mpmath.mpf(out[1]) - v) / v)))
    assert worst < eps / 16, (worst, eps)
    xs = np.exp(rng.uniform(-700, 700, 100000))
    assert all(H.core_log_cr(x) == H.core_log_cr_full(x) for x in xs)
    ys = rng.uniform(-700, 700, 100000)
    assert all(H.core_exp_cr(y) == H.core_exp_cr_full(y) for y in ys)


def _core_rpoly(H, op):
    op = np.ascontiguousarray(op, dtype=np.float64)
    zr, zi = np.zeros(8), np.zeros(8)
    H.core_rpoly.restype = ctypes.c_int
    k = H.core_rpoly(op.ctypes.data_as(dp), len(op) - 1, zr.ctypes.data_as(dp), zi.ctypes.data_as(dp))
    return zr[:max(k, 0)], zi[:max(k, 0)], k


@pytest.mark.skipif(not oracle.ref_rpoly_available(), reason="the reference's root finder (oracle/_ref) is not built")
def test_core_rpoly_restatement_matches_reference(core_harness, core_harness_libm):
    """Random, clustered, wide-range and sparse polynomials of degree 3..6: the restated solver returns
    the reference's roots bit for bit; with the C library's log/exp in place of the correctly rounded
    ones (the test-only variant) it does so on every polynomial, so the rare exception of the product
    build is the C library's misrounding of log/exp alone."""
    rng = np.random.default_rng(5)
    tot = bad = bad_libm = 0
    for trial in range(8000):
        n = int(rng.integers(3, 7))
        kind = trial % 4
        if kind == 0:
            op = rng.normal(size=n + 1)
        elif kind == 1:
            r = rng.uniform(-0.2, 1.2, size=n)
            r[1] = r[0] + rng.uniform(-1e-6, 1e-6)
            op = np.poly(r) * rng.uniform(0.1, 10)
        elif kind == 2:
            op = rng.normal(size=n + 1) * 10.0 ** rng.uniform(-30, 5, size=n + 1)
        else:
            op = rng.normal(size=n + 1)
            op[rng.integers(1, n + 1)] = 0.0
        op = op / np.abs(op).max()
        if op[0] == 0:
            continue
        a, b, c = oracle.ref_rpoly(op), _core_rpoly(core_harness, op), _core_rpoly(core_harness_libm, op)
        tot += 1
        same = lambda x: a[2] == x[2] and np.array_equal(a[0], x[0]) and np.array_equal(a[1], x[1])
        bad += not same(b)
        bad_libm += not same(c)
    print("rpoly restatement: mismatches", bad, "of", tot, "; with the C library's log/exp:", bad_libm)
    assert bad <= tot // 2000
    assert bad_libm == 0


@pytest.mark.skipif(not oracle.ref_rpoly_available(), reason="the reference's root finder (oracle/_ref) is not built")
@pytest.mark.parametrize("scene", ["converging", "cloth64"])
def test_core_reference_mode_vs_oracle(core_harness, core_harness_libm, scene):
    """Reference-root mode of the host-built core against the oracle running the reference's solver."""
    from tests.helpers import compare_reference, libm_core

    s = {"converging": scenes.converging_spheres, "cloth64": lambda: scenes.cloth(64, drop=0.05, noise=0.01)}[scene]()
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
    oracle.use_rootfinder("rpoly")
    core_harness.core_set_ref_roots(1)
    try:
        for keys, is_ee in ((vf, False), (ee, True)):
            rec = oracle.narrowphase(h, keys, s["eta"], is_ee)
            hit, t, sub = core_narrowphase(core_harness, s["qstart"], s["qend"], keys, is_ee, s["eta"])
            rep = compare_reference(scene, hit, t, rec,
                                    libm=libm_core(core_harness_libm, s["qstart"], s["qend"], keys, is_ee, s["eta"]))
            print(rep)
            assert not rep["hard"] and not rep["t_bad"], rep
            # every TOI that is not the reference's bit for bit is explained by the C library's log/exp
            assert rep["n_not_bit_identical"] == rep["n_libm_rounding"], rep
            both = hit & rec["hit"].astype(bool)
            assert np.array_equal(sub[both], rec["subtest"][both])
    finally:
        core_harness.core_set_ref_roots(0)
        oracle.use_default_rootfinder()


# ------------------------------------------------------------------------------------------------
# Separating-plane narrowphase (SeparatingPlaneNarrowPhase.cpp:9-278)
# ------------------------------------------------------------------------------------------------
def _sampled_min_distance(q0, q1, key, is_ee, samples):
    """Smallest distance between the stencil's primitives over `samples` times of straight motion."""
    best = np.inf
    for t in np.linspace(0.0, 1.0, samples):
        P = (1 - t) * q0[key] + t * q1[key]
        vec, _ = oracle.distance(3 if is_ee else 2, P.reshape(-1))
        best = min(best, float(np.sqrt(vec @ vec)))
    return best


def test_oracle_separating_plane_pinned_by_sampled_distance():
    """The oracle's separating-plane restatement against what the method guarantees: a stencil it clears
    stays at least eta apart (each side keeps 0.5 eta from a moving plane), and one it reports comes
    within eta.  Checked on sampled minimum distances of straight-line motion (a flipped plane normal,
    a wrong bisection bound or a dropped sub-interval breaks one side)."""
    s = scenes.cloth(24, drop=0.05, noise=0.01)
    eta = 0.01
    h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    vf, ee = oracle.candidates(s["faces"], h, eta, method="tree")
    rng = np.random.default_rng(3)
    for keys, is_ee in ((vf, False), (ee, True)):
        st, _, _ = oracle.separating_plane(h, keys, eta, is_ee)
        assert not (st == 2).any() and (st == 1).sum() >= 10
        for i in np.where(st == 1)[0][:40]:
            assert _sampled_min_distance(s["qstart"], s["qend"], keys[i], is_ee, 801) < 1.02 * eta, (is_ee, i)
        for i in rng.choice(np.where(st == 0)[0], 60, replace=False):
            assert _sampled_min_distance(s["qstart"], s["qend"], keys[i], is_ee, 201) > 0.98 * eta, (is_ee, i)


@pytest.mark.parametrize("hist_kind", ["two_entry", "three_entry"])
def test_core_separating_plane_equals_oracle(core_harness, hist_kind):
    """The core's explicit-stack restatement equals the oracle's recursive one, status for status
    (undecided included), in both root modes."""
    from tests.helpers import core_separating_plane, three_entry_history

    if hist_kind == "two_entry":
        s = scenes.cloth(48, drop=0.05, noise=0.01)
        h = oracle.Hist.two_entry(s["qstart"], s["qend"])
    else:
        s = scenes.converging_spheres()
        h = oracle.Hist(*three_entry_history(s))
    vf, ee = oracle.candidates(s["faces"], h, s["eta"], method="tree")
    for mode in ((0, "aberth"), (1, "rpoly")) if oracle.ref_rpoly_available() else ((0, "aberth"),):
        core_harness.core_set_ref_roots(mode[0])
        oracle.use_rootfinder(mode[1])
        try:
            for keys, is_ee in ((vf, False), (ee, True)):
                st, margin, depth = oracle.separating_plane(h, keys, s["eta"], is_ee)
                got = core_separating_plane(core_harness, h, keys, is_ee, s["eta"])
                diff = np.where(st != got)[0]
                # isolating roots vs the oracle's exact (Aberth) roots: only borderline decisions may differ
                assert all(margin[i] < 1e-12 for i in diff) if mode[0] == 0 else len(diff) == 0, (mode, is_ee, diff[:5])
                print(hist_kind, mode, is_ee, np.bincount(st, minlength=3), "max depth", depth.max())
                assert (st == 1).sum() > 0
        finally:
            core_harness.core_set_ref_roots(0)
            oracle.use_default_rootfinder()

