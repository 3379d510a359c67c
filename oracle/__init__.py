"""TEST INFRASTRUCTURE ONLY -- the CPU oracle for the swept-primitive CCD path.

A ctypes front end for ``oracle/liborc.so`` (``oracle/ccd_oracle.cpp``), the CPU restatement of
evouga/collisiondetection's KDOPBroadPhase -> CTCDNarrowPhase -> CTCD path.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this package, and only
as the checker (or the timed CPU baseline).  The product (``collisiondetection_amd``) never imports it.

Parity status of each function (what pins it) is listed in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_REF_RPOLY_PATH = os.path.join(_HERE, "_ref", "librpoly_ref.so")
REFERENCE_DIR = "/root/reference"

_lib = None
_ref = None

# Stage counters reported by orc_narrowphase (order of the Stage enum in ccd_oracle.cpp).
STAGES = ("vf_call", "vf_plane", "vf_dist", "ee_dist", "ee_par", "ee_bary", "ve_call", "ve_quad", "ve_quart",
          "vv", "quad_root", "isolate")

RECORD_DTYPE = np.dtype([("hit", np.int32), ("subtest", np.int32), ("window", np.int32),
                         ("pad", np.int32), ("t", np.float64), ("margin", np.float64)])

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_up = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> None:
    """Compile liborc.so (and the reference rpoly checker when the reference tree is present)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s", "liborc.so"], check=True,
                       stdout=subprocess.DEVNULL)
    if os.path.isdir(REFERENCE_DIR) and (force or not os.path.exists(_REF_RPOLY_PATH)):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s", "ref"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_set_dot_order.argtypes = [ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_rootfinder.argtypes = [ctypes.c_void_p]
        L.orc_aberth_roots.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.orc_aberth_roots.restype = ctypes.c_int
        L.orc_find_roots.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.orc_find_roots.restype = ctypes.c_int
        L.orc_kdop_axes.argtypes = [_dp]
        L.orc_stitch.argtypes = [_lp, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int64, _dp, _dp]
        L.orc_stitch.restype = ctypes.c_int64
        L.orc_sampled_min_distance.argtypes = [_dp, _dp, _ip, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_get_quad_roots.argtypes = [ctypes.c_double] * 3 + [_dp, _dp]
        L.orc_get_quad_roots.restype = ctypes.c_int
        L.orc_find_intervals.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.orc_find_intervals.restype = ctypes.c_int
        for nm in ("orc_vertex_face_ctcd", "orc_edge_edge_ctcd", "orc_vertex_edge_ctcd",
                   "orc_vertex_vertex_ctcd"):
            f = getattr(L, nm)
            f.argtypes = [_dp, ctypes.c_double, _dp, _dp]
            f.restype = ctypes.c_int
        L.orc_leaf_dops.argtypes = [_ip, ctypes.c_int64, _lp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.orc_candidates.argtypes = [_ip, ctypes.c_int64, _lp, _dp, _dp, ctypes.c_double, _up, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
        L.orc_candidates.restype = ctypes.c_void_p
        L.orc_result_count.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_result_count.restype = ctypes.c_int64
        L.orc_result_copy.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.orc_result_free.argtypes = [ctypes.c_void_p]
        L.orc_narrowphase.argtypes = [_lp, _dp, _dp, _ip, _dp, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, _lp]
        L.orc_stage_count.restype = ctypes.c_int
        L.orc_distance.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]
        L.orc_distance.restype = ctypes.c_int
        L.orc_mesh_self_distance.argtypes = [_dp, ctypes.c_int64, _ip, ctypes.c_int64, _up, _lp, _lp]
        L.orc_mesh_self_distance.restype = ctypes.c_double
        L.orc_minimum_gap.argtypes = [_lp, _dp, _dp, ctypes.c_int64]
        L.orc_minimum_gap.restype = ctypes.c_double
        L.orc_separating_plane.argtypes = [_lp, _dp, _dp, ctypes.c_int64, _ip, _dp, ctypes.c_int64, ctypes.c_int,
                                           _ip, _dp, _ip]
        assert L.orc_stage_count() == len(STAGES)
        _lib = L
    return _lib


def ref_rpoly_available() -> bool:
    return os.path.exists(_REF_RPOLY_PATH)


def _ref_lib():
    global _ref
    if _ref is None:
        if not ref_rpoly_available():
            raise FileNotFoundError(_REF_RPOLY_PATH)
        _ref = ctypes.CDLL(_REF_RPOLY_PATH)
        _ref.ref_rpoly.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        _ref.ref_rpoly.restype = ctypes.c_int
    return _ref


_ROOTFN = ctypes.CFUNCTYPE(ctypes.c_int, _dp, ctypes.c_int, _dp, _dp)
_mp_callback = None


def _mp_roots(op, degree, zr, zi):
    """Roots of the double polynomial op[0..degree] at 60 significant digits (mpmath), real parts and
    imaginary parts rounded to double; exact zero roots first, as rpoly reports them."""
    import mpmath as mp

    c = [op[i] for i in range(degree + 1)]
    if c[0] == 0.0:
        return -1
    nz = 0
    while degree > 0 and c[degree] == 0.0:
        zr[nz] = 0.0
        zi[nz] = 0.0
        nz += 1
        degree -= 1
        c = c[:-1]
    if degree == 0:
        return nz
    with mp.workdps(60):
        coeffs = [mp.mpf(x) for x in c]
        roots = None
        for steps in (200, 2000, 20000):
            try:
                roots = mp.polyroots(coeffs, maxsteps=steps, extraprec=400)
                break
            except mp.NoConvergence:
                continue
        if roots is None:
            return nz
        for k, z in enumerate(roots):
            zr[nz + k] = float(mp.re(z))
            zi[nz + k] = float(mp.im(z))
    return nz + degree


def use_rootfinder(which: str) -> None:
    """Select the oracle's root finder: 'aberth' (built in), 'rpoly' (the reference's, from _ref) or
    'mpmath' (the double polynomial's roots to 60 digits -- slow; the arbiter for disagreements)."""
    global _mp_callback
    if which == "aberth":
        lib().orc_set_rootfinder(None)
    elif which == "rpoly":
        fn = ctypes.cast(_ref_lib().ref_rpoly, ctypes.c_void_p).value
        lib().orc_set_rootfinder(fn)
    elif which == "mpmath":
        if _mp_callback is None:
            _mp_callback = _ROOTFN(_mp_roots)
        lib().orc_set_rootfinder(ctypes.cast(_mp_callback, ctypes.c_void_p).value)
    else:
        raise ValueError(which)


def use_default_rootfinder() -> str:
    """The reference's own rpoly when oracle/_ref holds it, else the built-in Aberth iteration."""
    which = "rpoly" if ref_rpoly_available() else "aberth"
    use_rootfinder(which)
    return which


def set_dot_order(order: int) -> None:
    lib().orc_set_dot_order(int(order))


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _l(a):
    return a.ctypes.data_as(_lp)


# ------------------------------------------------------------------------------------------------
# Roots and intervals
# ------------------------------------------------------------------------------------------------
def aberth_roots(op):
    op = np.ascontiguousarray(op, dtype=np.float64)
    n = len(op) - 1
    zr = np.zeros(8)
    zi = np.zeros(8)
    k = lib().orc_aberth_roots(_d(op), n, _d(zr), _d(zi))
    return zr[:max(k, 0)], zi[:max(k, 0)], k


def ref_rpoly(op):
    op = np.array(op, dtype=np.float64)
    n = len(op) - 1
    zr = np.zeros(8)
    zi = np.zeros(8)
    k = _ref_lib().ref_rpoly(_d(op), n, _d(zr), _d(zi))
    return zr[:max(k, 0)], zi[:max(k, 0)], k


def get_quad_roots(a, b, c):
    t0 = ctypes.c_double(0.0)
    t1 = ctypes.c_double(0.0)
    k = lib().orc_get_quad_roots(a, b, c, ctypes.byref(t0), ctypes.byref(t1))
    return k, t0.value, t1.value


def find_intervals(op, pos: bool):
    """CTCD::findIntervals on a copy of op -> (list of (l,u), normalised op, margin)."""
    op = np.ascontiguousarray(op, dtype=np.float64)
    n = len(op) - 1
    out = np.zeros(32)
    nop = np.zeros(n + 1)
    m = ctypes.c_double(0.0)
    k = lib().orc_find_intervals(_d(op), n, int(pos), _d(out), 16, _d(nop), ctypes.byref(m))
    return [(out[2 * i], out[2 * i + 1]) for i in range(k)], nop, m.value


# ------------------------------------------------------------------------------------------------
# Primitives: q holds the start positions then the end positions, as in the CTCD statics.
# ------------------------------------------------------------------------------------------------
def _prim(name, q, eta):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    t = ctypes.c_double(-1.0)
    m = ctypes.c_double(0.0)
    r = getattr(lib(), name)(_d(q), float(eta), ctypes.byref(t), ctypes.byref(m))
    return bool(r), (t.value if r else None), m.value


def vertex_face_ctcd(q, eta):
    return _prim("orc_vertex_face_ctcd", q, eta)


def edge_edge_ctcd(q, eta):
    return _prim("orc_edge_edge_ctcd", q, eta)


def vertex_edge_ctcd(q, eta):
    return _prim("orc_vertex_edge_ctcd", q, eta)


def vertex_vertex_ctcd(q, eta):
    return _prim("orc_vertex_vertex_ctcd", q, eta)


# ------------------------------------------------------------------------------------------------
# Histories
# ------------------------------------------------------------------------------------------------
class Hist:
    """CSR history: vertex v owns entries off[v]..off[v+1]-1 (times ht, positions hp[3*])."""

    def __init__(self, off, ht, hp):
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.ht = np.ascontiguousarray(ht, dtype=np.float64)
        self.hp = np.ascontiguousarray(hp, dtype=np.float64).reshape(-1)

    @staticmethod
    def two_entry(qstart, qend):
        """History(qstart) + finishHistory(qend) (History.cpp:8-39): entries at t=0 and t=1."""
        qs = np.asarray(qstart, dtype=np.float64).reshape(-1, 3)
        qe = np.asarray(qend, dtype=np.float64).reshape(-1, 3)
        V = qs.shape[0]
        off = np.arange(0, 2 * V + 1, 2, dtype=np.int64)
        ht = np.tile(np.array([0.0, 1.0]), V)
        hp = np.stack([qs, qe], axis=1).reshape(-1)
        return Hist(off, ht, hp)

    @property
    def nverts(self):
        return len(self.off) - 1


def stitch(hist: Hist, verts):
    """History::stitchCommonHistory (History.cpp:98-140) for up to 4 vertices: (times (n,), positions
    (n, len(verts), 3))."""
    v = np.ascontiguousarray(verts, dtype=np.int32)
    nv = len(v)
    cap = int(sum(hist.off[x + 1] - hist.off[x] for x in v)) + 2
    t = np.zeros(cap)
    p = np.zeros((cap, nv, 3))
    n = lib().orc_stitch(_l(hist.off), _d(hist.ht), _d(hist.hp), _i(v), nv, cap, _d(t), _d(p))
    assert n <= cap
    return t[:n], p[:n]


def sampled_min_distance(qstart, qend, keys, is_ee: bool, samples: int = 401):
    """Per stencil, the smallest vertex-face / edge-edge distance over `samples` equally spaced times of
    the straight motion qstart -> qend (Distance.h's closest points)."""
    keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 4)
    qstart = np.ascontiguousarray(qstart, dtype=np.float64)
    qend = np.ascontiguousarray(qend, dtype=np.float64)
    out = np.zeros(len(keys))
    lib().orc_sampled_min_distance(_d(qstart), _d(qend), _i(keys), len(keys), int(is_ee), int(samples), _d(out))
    return out


def kdop_axes():
    out = np.zeros(39)
    lib().orc_kdop_axes(_d(out))
    return out.reshape(13, 3)


def leaf_dops(faces, hist: Hist, eta, aabb=False):
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
    F = len(faces) // 3
    mn = np.zeros((F, 13))
    mx = np.zeros((F, 13))
    lib().orc_leaf_dops(_i(faces), F, _l(hist.off), _d(hist.ht), _d(hist.hp), float(eta), int(aabb), _d(mn), _d(mx))
    return mn, mx


METHODS = {"brute": 0, "tree_set": 1, "tree": 2, "tree_parallel": 3}


def set_threads(n: int) -> None:
    """Worker threads for the all-core CPU baseline (the tree_parallel traversal and the narrowphase);
    1 (the default) is the reference's single thread."""
    lib().orc_set_threads(int(n))


def candidates(faces, hist: Hist, eta, fixed=None, method="brute", aabb=False, want_pairs=False):
    """Candidate stencil sets (sorted, unique) as int32 arrays of shape (n, 4): (vf, ee[, pairs])."""
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
    F = len(faces) // 3
    fx = None
    if fixed is not None:
        fx_arr = np.zeros(hist.nverts, dtype=np.uint8)
        fx_arr[np.asarray(list(fixed), dtype=np.int64)] = 1
        fx = fx_arr.ctypes.data_as(_up)
    L = lib()
    h = L.orc_candidates(_i(faces), F, _l(hist.off), _d(hist.ht), _d(hist.hp), float(eta), fx,
                         METHODS[method], int(aabb), int(want_pairs))
    try:
        out = []
        for which in (0, 1):
            n = L.orc_result_count(h, which)
            a = np.zeros((n, 4), dtype=np.int32)
            if n:
                L.orc_result_copy(h, which, _i(a))
            out.append(a)
        if want_pairs:
            n = L.orc_result_count(h, 2)
            p = np.zeros((n, 2), dtype=np.int32)
            if n:
                L.orc_result_copy(h, 2, _i(p))
            out.append(p)
    finally:
        L.orc_result_free(h)
    return tuple(out)


def narrowphase(hist: Hist, keys, eta, is_ee: bool, with_counts=False):
    """CTCDNarrowPhase per stencil -> structured records (hit, subtest, window, t, margin)."""
    keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 4)
    n = keys.shape[0]
    eta_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(eta, dtype=np.float64), (n,)))
    rec = np.zeros(n, dtype=RECORD_DTYPE)
    counts = np.zeros(len(STAGES), dtype=np.int64)
    if n:
        lib().orc_narrowphase(_l(hist.off), _d(hist.ht), _d(hist.hp), _i(keys), _d(eta_arr), n, int(is_ee),
                              rec.ctypes.data_as(ctypes.c_void_p), _l(counts))
    if with_counts:
        return rec, dict(zip(STAGES, counts.tolist()))
    return rec


def minimum_gap(hist: Hist) -> float:
    """History::computeMinimumGap (History.cpp:78-93)."""
    return lib().orc_minimum_gap(_l(hist.off), _d(hist.ht), _d(hist.hp), hist.nverts)


def separating_plane(hist: Hist, keys, eta, is_ee: bool):
    """SeparatingPlaneNarrowPhase per stencil (SeparatingPlaneNarrowPhase.cpp:9-278), recursive as the
    reference is -> (status array: 1 collision, 0 none, 2 undecided -- the reference's recursion would not
    terminate; decision margin of the primitive tests; deepest recursion level)."""
    keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 4)
    n = keys.shape[0]
    eta_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(eta, dtype=np.float64), (n,)))
    hit = np.zeros(n, dtype=np.int32)
    margin = np.zeros(n, dtype=np.float64)
    depth = np.zeros(n, dtype=np.int32)
    if n:
        lib().orc_separating_plane(_l(hist.off), _d(hist.ht), _d(hist.hp), hist.nverts, _i(keys), _d(eta_arr), n,
                                   int(is_ee), _i(hit), _d(margin), _i(depth))
    return hit, margin, depth


# ------------------------------------------------------------------------------------------------
# Proximity (include/Distance.h, src/Distance.cpp)
# ------------------------------------------------------------------------------------------------
def distance(kind: int, q, eta: float = 0.0):
    """kind 0/1: (flag,) ; kind 2/3: (vec[3], bary[4])."""
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    vec = np.zeros(3)
    bary = np.zeros(4)
    r = lib().orc_distance(int(kind), _d(q), float(eta), _d(vec), _d(bary))
    if kind < 2:
        return bool(r)
    return vec, bary


def mesh_self_distance(verts, faces, fixed=None):
    """Distance::meshSelfDistance -> (distance, n_vf_candidates, n_ee_candidates)."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
    V, F = len(v) // 3, len(f) // 3
    fx = None
    if fixed is not None:
        fx_arr = np.zeros(V, dtype=np.uint8)
        fx_arr[np.asarray(list(fixed), dtype=np.int64)] = 1
        fx = fx_arr.ctypes.data_as(_up)
    nv = ctypes.c_int64(0)
    ne = ctypes.c_int64(0)
    d = lib().orc_mesh_self_distance(_d(v), V, _i(f), F, fx, ctypes.byref(nv), ctypes.byref(ne))
    return d, nv.value, ne.value


# ------------------------------------------------------------------------------------------------
# Velocity filter (include/VelocityFilter.h, src/VelocityFilter.cpp, src/ActiveLayers.cpp,
# src/PenaltyGroup.cpp, src/PenaltyPotential.cpp)
# ------------------------------------------------------------------------------------------------
VF_STATS = ("iterations", "layers", "firings", "history_entries", "vf_candidates", "ee_candidates",
            "vf_collisions", "ee_collisions", "undecided")


def velocity_filter(qstart, qend, faces, inv_masses, outer_eta, inner_eta, base_stiffness=1e-4,
                    base_substep=1e-4, cor=0.1, max_rollbacks=1000, want_history=False):
    """VelocityFilter::velocityFilter -> (result, qend_out, stats, closest[, hist]).

    result is the reference's return value (iterations, -1 dangling vertices, -2 starting state
    intersects, -3 rollbacks exceeded), or -100 where the separating-plane recursion never ends.
    hist (want_history, on success) is the final iteration's history as a Hist."""
    L = lib()
    if not hasattr(L, "_vf_bound"):
        L.orc_velocity_filter.argtypes = [_dp, _dp, ctypes.c_int64, _ip, ctypes.c_int64, _dp, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, _lp, _dp, ctypes.POINTER(ctypes.c_void_p)]
        L.orc_velocity_filter.restype = ctypes.c_int
        L.orc_vhist_entries.argtypes = [ctypes.c_void_p]
        L.orc_vhist_entries.restype = ctypes.c_int64
        L.orc_vhist_copy.argtypes = [ctypes.c_void_p, _lp, _dp, _dp]
        L.orc_vhist_free.argtypes = [ctypes.c_void_p]
        L._vf_bound = True
    qs = np.ascontiguousarray(qstart, dtype=np.float64).reshape(-1)
    qe = np.array(qend, dtype=np.float64).reshape(-1)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
    im = np.ascontiguousarray(inv_masses, dtype=np.float64).reshape(-1)
    V, F = len(qs) // 3, len(f) // 3
    assert len(qe) == 3 * V and len(im) == 3 * V
    stats = np.zeros(len(VF_STATS), dtype=np.int64)
    closest = ctypes.c_double(0.0)
    h = ctypes.c_void_p(None)
    r = L.orc_velocity_filter(_d(qs), _d(qe), V, _i(f), F, _d(im), float(outer_eta), float(inner_eta),
                              float(base_stiffness), float(base_substep), float(cor), int(max_rollbacks), _l(stats),
                              ctypes.byref(closest), ctypes.byref(h) if want_history else None)
    st = dict(zip(VF_STATS, (int(x) for x in stats)))
    if not want_history:
        return r, qe, st, closest.value
    hist = None
    if h.value:
        E = L.orc_vhist_entries(h)
        off = np.zeros(V + 1, dtype=np.int64)
        ht = np.zeros(E)
        hp = np.zeros(3 * E)
        L.orc_vhist_copy(h, _l(off), _d(ht), _d(hp))
        L.orc_vhist_free(h)
        hist = Hist(off, ht, hp)
    return r, qe, st, closest.value, hist
