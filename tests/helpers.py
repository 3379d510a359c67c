"""Shared helpers for the test suite (not collected)."""
import ctypes

import numpy as np

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int32)


def core_narrowphase(H, qs, qe, keys, is_ee, eta):
    keys = np.ascontiguousarray(keys, np.int32).reshape(-1, 4)
    n = len(keys)
    hit = np.zeros(n, np.int32)
    t = np.zeros(n)
    sub = np.zeros(n, np.int32)
    qs = np.ascontiguousarray(qs, dtype=np.float64)
    qe = np.ascontiguousarray(qe, dtype=np.float64)
    H.core_narrowphase(qs.ctypes.data_as(dp), qe.ctypes.data_as(dp), keys.ctypes.data_as(ip), ctypes.c_int64(n),
                       int(is_ee), ctypes.c_double(eta), hit.ctypes.data_as(ip), t.ctypes.data_as(dp),
                       sub.ctypes.data_as(ip))
    return hit.astype(bool), t, sub


def core_ee_two_pass(H, qs, qe, keys, eta):
    """The host-built core's two-pass edge-edge schedule: (hit, t, sub, deferred) per stencil."""
    keys = np.ascontiguousarray(keys, np.int32).reshape(-1, 4)
    n = len(keys)
    hit = np.zeros(n, np.int32)
    t = np.zeros(n)
    sub = np.zeros(n, np.int32)
    dfr = np.zeros(n, np.int32)
    qs = np.ascontiguousarray(qs, dtype=np.float64)
    qe = np.ascontiguousarray(qe, dtype=np.float64)
    H.core_ee_two_pass.restype = ctypes.c_int64
    H.core_ee_two_pass(qs.ctypes.data_as(dp), qe.ctypes.data_as(dp), keys.ctypes.data_as(ip), ctypes.c_int64(n),
                       ctypes.c_double(eta), hit.ctypes.data_as(ip), t.ctypes.data_as(dp), sub.ctypes.data_as(ip),
                       dfr.ctypes.data_as(ip))
    return hit.astype(bool), t, sub, dfr.astype(bool)


def core_separating_plane(H, hist, keys, is_ee, eta):
    """The host-built core's separating-plane narrowphase: statuses 1 / 0 / 2 (undecided)."""
    keys = np.ascontiguousarray(keys, np.int32).reshape(-1, 4)
    n = len(keys)
    st = np.zeros(n, np.int32)
    e = np.ascontiguousarray(np.broadcast_to(np.asarray(eta, dtype=np.float64), (n,)))
    lp = ctypes.POINTER(ctypes.c_int64)
    H.core_separating_plane(hist.off.ctypes.data_as(lp), hist.ht.ctypes.data_as(dp), hist.hp.ctypes.data_as(dp),
                            ctypes.c_int64(hist.nverts), keys.ctypes.data_as(ip), e.ctypes.data_as(dp),
                            ctypes.c_int64(n), int(is_ee), st.ctypes.data_as(ip))
    return st


def three_entry_history(s, seed=9, wobble=0.02):
    """A bent path through a per-vertex midpoint: entries at 0, tmid ~ U(0.2, 0.8) and 1."""
    rng = np.random.default_rng(seed)
    V = len(s["qstart"])
    qm = 0.5 * (s["qstart"] + s["qend"]) + rng.normal(0, wobble, size=(V, 3))
    tmid = rng.uniform(0.2, 0.8, size=V)
    off = np.arange(0, 3 * V + 1, 3, dtype=np.int64)
    ht = np.stack([np.zeros(V), tmid, np.ones(V)], axis=1).reshape(-1)
    hp = np.stack([s["qstart"], qm, s["qend"]], axis=1).reshape(-1)
    return off, ht, hp


def core_primitive(H, kind, q, eta):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    t = ctypes.c_double(0.0)
    r = H.core_primitive(int(kind), q.ctypes.data_as(dp), ctypes.c_double(eta), ctypes.byref(t))
    return bool(r), (t.value if r else None)


def core_find_intervals(H, op, pos):
    op = np.ascontiguousarray(op, dtype=np.float64)
    out = np.zeros(32)
    k = H.core_find_intervals(op.ctypes.data_as(dp), len(op) - 1, int(pos), out.ctypes.data_as(dp))
    return [(out[2 * i], out[2 * i + 1]) for i in range(k)]


def core_roots01(H, c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    r = np.zeros(8)
    k = H.core_roots01(c.ctypes.data_as(dp), len(c) - 1, r.ctypes.data_as(dp))
    return r[:k]


def compare_results(name, hit_a, t_a, rec, tol_t=1e-9, margin_tol=1e-12, arbiter=None):
    """Compare a product result (hit, t) with oracle records.

    Flag mismatches whose oracle margin is below margin_tol are listed as borderline.  With an arbiter
    (a function idx -> oracle records recomputed with exact roots, oracle root finder 'mpmath', or a
    list of such functions tried in turn), every other disagreement -- a flag, or a time of impact
    more than tol_t apart -- is re-decided against the exact-root answer: if the product agrees with
    it, the disagreement is the reference root finder's error and is listed under ref_root_error;
    otherwise it is hard."""
    hit_o = rec["hit"].astype(bool)
    mism = np.where(hit_a != hit_o)[0]
    border = mism[rec["margin"][mism] < margin_tol]
    hard = mism[rec["margin"][mism] >= margin_tol]
    both = hit_a & hit_o
    tdiff = np.abs(t_a[both] - rec["t"][both])
    idx_both = np.where(both)[0]
    t_bad = idx_both[tdiff > tol_t]
    t_bad_hard = t_bad[rec["margin"][t_bad] >= margin_tol]
    ref_err = []
    arbiters = [] if arbiter is None else (list(arbiter) if isinstance(arbiter, (list, tuple)) else [arbiter])
    if arbiters and (len(hard) or len(t_bad_hard)):
        cand = np.unique(np.concatenate([hard, t_bad_hard])).astype(np.int64)
        hard_set = set(hard.tolist())
        keep_hard, keep_t = [], []
        # Each arbiter re-decides what the previous ones left open; the last one's verdict stands.
        for level, arb in enumerate(arbiters):
            ex = arb(cand)
            left = []
            for j, i in enumerate(cand):
                eh = bool(ex["hit"][j])
                ok = (eh == bool(hit_a[i])) and (not eh or abs(ex["t"][j] - t_a[i]) <= tol_t)
                if ok:
                    ref_err.append((int(i), float(rec["t"][i]), float(ex["t"][j]), float(t_a[i])))
                elif ex["margin"][j] < margin_tol:
                    pass
                elif level + 1 < len(arbiters):
                    left.append(int(i))
                elif i in hard_set:
                    keep_hard.append(int(i))
                else:
                    keep_t.append(int(i))
            cand = np.array(left, dtype=np.int64)
            if not len(cand):
                break
        hard = np.array(keep_hard, dtype=np.int64)
        t_bad_hard = np.array(keep_t, dtype=np.int64)
    return {"name": name, "n": len(hit_a), "hits": int(hit_o.sum()), "mismatch": len(mism),
            "borderline": border.tolist()[:20], "hard": hard.tolist()[:20],
            "max_tdiff": float(tdiff.max()) if len(tdiff) else 0.0, "t_bad": t_bad.tolist()[:20],
            "t_bad_hard": t_bad_hard.tolist()[:20], "ref_root_error": ref_err[:20], "n_ref_root_error": len(ref_err)}


def aberth_arbiter(hist, keys, eta, is_ee):
    """idx -> oracle records for keys[idx] computed with the oracle's long-double Aberth roots (accurate
    to ~1e-16 on these polynomials, and fast): a first arbiter, backed by mpmath_arbiter."""
    import oracle

    def run(idx):
        eta_arr = np.broadcast_to(np.asarray(eta, dtype=np.float64), (len(keys),))
        try:
            oracle.use_rootfinder("aberth")
            return oracle.narrowphase(hist, keys[idx], eta_arr[idx], is_ee)
        finally:
            oracle.use_default_rootfinder()

    return run


def exact_arbiters(hist, keys, eta, is_ee):
    """Aberth first, 60-digit mpmath for whatever Aberth leaves in dispute."""
    return [aberth_arbiter(hist, keys, eta, is_ee), mpmath_arbiter(hist, keys, eta, is_ee)]


def mpmath_arbiter(hist, keys, eta, is_ee):
    """idx -> oracle records for keys[idx] computed with exact (60-digit) roots."""
    import oracle

    def run(idx):
        eta_arr = np.broadcast_to(np.asarray(eta, dtype=np.float64), (len(keys),))
        try:
            oracle.use_rootfinder("mpmath")
            return oracle.narrowphase(hist, keys[idx], eta_arr[idx], is_ee)
        finally:
            oracle.use_default_rootfinder()

    return run


def core_distance(H, kind, q, eta=0.0):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    vec = np.zeros(3)
    bary = np.zeros(4)
    r = H.core_distance(int(kind), q.ctypes.data_as(dp), ctypes.c_double(eta), vec.ctypes.data_as(dp),
                        bary.ctypes.data_as(dp))
    if kind < 2:
        return bool(r)
    return vec, bary
