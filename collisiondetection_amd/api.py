"""Python front end of the MI355X CCD engine.

Two layers:
  * array functions (``primitives``, ``narrowphase``, ``narrowphase_history``, ``kdop_candidates``,
    ``detect``) that take numpy arrays or torch tensors, keep data in HBM and call the C ABI;
  * classes mirroring the reference's plugin interface for this path -- ``Mesh``, ``History``,
    ``KDOPBroadPhase``, ``CTCDNarrowPhase`` and ``CTCD`` (src/Mesh.h, src/History.h,
    src/RetrospectiveDetection.h, src/KDOPBroadPhase.h, src/CTCDNarrowPhase.h, include/CTCD.h) -- with the
    same names, argument meaning and set-valued results.

Device memory and streams come from torch (HIP).  Every computation runs in libccd_amd.so on the
GPU; nothing here computes a result on the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N

_ctx_cache = {}


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("collisiondetection_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch


class Context:
    """Owns a ccd_ctx (stream + grow-only device arena) on one device."""

    def __init__(self, device: int = 0):
        torch = _torch()
        self.device = device
        self.torch_device = torch.device("cuda", device)
        self._p = ctypes.c_void_p()
        rc = N.lib().ccd_create(ctypes.byref(self._p), device)
        if rc != N.CCD_OK:
            raise N.CCDError(rc, "ccd_create failed")

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                N.lib().ccd_destroy(p)
            except Exception:
                pass
            self._p = None

    @property
    def ptr(self):
        return self._p

    def bind_stream(self):
        torch = _torch()
        s = torch.cuda.current_stream(self.torch_device)
        N.lib().ccd_set_stream(self._p, ctypes.c_void_p(s.cuda_stream))

    def check(self, rc: int):
        if rc != N.CCD_OK:
            msg = N.lib().ccd_last_error(self._p)
            raise N.CCDError(rc, msg.decode() if msg else "")

    def device_bytes(self) -> int:
        return int(N.lib().ccd_device_bytes(self._p))


def default_context(device: int = 0) -> Context:
    if device not in _ctx_cache:
        _ctx_cache[device] = Context(device)
    return _ctx_cache[device]


def to_device(x, dtype, ctx: Context):
    """A contiguous torch tensor of `dtype` on ctx's device (no copy if it already is one)."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(x))
    if t.dtype != dtype:
        t = t.to(dtype)
    if t.device != ctx.torch_device:
        t = t.to(ctx.torch_device, non_blocking=False)
    return t.contiguous()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


# ------------------------------------------------------------------------------------------------
# Array layer
# ------------------------------------------------------------------------------------------------
_PRIM_WIDTH = {N.PRIM_VERTEX_FACE: 24, N.PRIM_EDGE_EDGE: 24, N.PRIM_VERTEX_EDGE: 18, N.PRIM_VERTEX_VERTEX: 12}


def primitives(kind: int, q, eta, ctx: Context | None = None, as_numpy: bool = True):
    """CTCD primitive tests on n items; q has shape (n, S) with S = 24/24/18/12 (starts, then ends)."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    S = _PRIM_WIDTH[kind]
    qd = to_device(q, torch.float64, ctx).reshape(-1)
    n = qd.numel() // S
    eta_arr = np.asarray(eta, dtype=np.float64)
    eta_d = None if eta_arr.ndim == 0 else to_device(eta_arr, torch.float64, ctx)
    hit = torch.empty(n, dtype=torch.uint8, device=ctx.torch_device)
    t = torch.empty(n, dtype=torch.float64, device=ctx.torch_device)
    ctx.check(N.lib().ccd_ctcd_primitives(ctx.ptr, kind, _ptr(qd), _ptr(eta_d), float(eta_arr.reshape(-1)[0]) if n else 0.0,
                                          n, _ptr(hit), _ptr(t)))
    if as_numpy:
        torch.cuda.synchronize(ctx.torch_device)
        return hit.cpu().numpy().astype(bool), t.cpu().numpy()
    return hit, t


def narrowphase(qstart, qend, vf_keys, ee_keys, eta, vf_eta=None, ee_eta=None, ctx: Context | None = None,
                as_numpy: bool = True):
    """CTCDNarrowPhase over stencil lists for two-entry histories.

    Returns dict(vf_hit, vf_t, vf_sub, ee_hit, ee_t, ee_sub)."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    qs = to_device(qstart, torch.float64, ctx).reshape(-1)
    qe = to_device(qend, torch.float64, ctx).reshape(-1)
    V = qs.numel() // 3
    vk = to_device(vf_keys, torch.int32, ctx).reshape(-1)
    ek = to_device(ee_keys, torch.int32, ctx).reshape(-1)
    nvf, nee = vk.numel() // 4, ek.numel() // 4
    ve = None if vf_eta is None else to_device(vf_eta, torch.float64, ctx)
    ee_ = None if ee_eta is None else to_device(ee_eta, torch.float64, ctx)
    dev = ctx.torch_device
    out = {
        "vf_hit": torch.empty(nvf, dtype=torch.uint8, device=dev),
        "vf_t": torch.empty(nvf, dtype=torch.float64, device=dev),
        "vf_sub": torch.empty(nvf, dtype=torch.uint8, device=dev),
        "ee_hit": torch.empty(nee, dtype=torch.uint8, device=dev),
        "ee_t": torch.empty(nee, dtype=torch.float64, device=dev),
        "ee_sub": torch.empty(nee, dtype=torch.uint8, device=dev),
    }
    ctx.check(N.lib().ccd_ctcd_narrowphase(
        ctx.ptr, _ptr(qs), _ptr(qe), V, _ptr(vk), _ptr(ve), nvf, _ptr(ek), _ptr(ee_), nee, float(eta),
        _ptr(out["vf_hit"]), _ptr(out["vf_t"]), _ptr(out["vf_sub"]), _ptr(out["ee_hit"]), _ptr(out["ee_t"]),
        _ptr(out["ee_sub"])))
    if as_numpy:
        torch.cuda.synchronize(dev)
        out = {k: v.cpu().numpy() for k, v in out.items()}
        out["vf_hit"] = out["vf_hit"].astype(bool)
        out["ee_hit"] = out["ee_hit"].astype(bool)
    return out


def narrowphase_history(hist_off, hist_t, hist_pos, vf_keys, ee_keys, eta, ctx: Context | None = None):
    """CTCDNarrowPhase over stencil lists for CSR multi-entry histories."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    off = to_device(hist_off, torch.int64, ctx)
    ht = to_device(hist_t, torch.float64, ctx)
    hp = to_device(hist_pos, torch.float64, ctx).reshape(-1)
    V = off.numel() - 1
    E = ht.numel()
    vk = to_device(vf_keys, torch.int32, ctx).reshape(-1)
    ek = to_device(ee_keys, torch.int32, ctx).reshape(-1)
    nvf, nee = vk.numel() // 4, ek.numel() // 4
    dev = ctx.torch_device
    o = {k: torch.empty(n, dtype=dt, device=dev) for k, n, dt in (
        ("vf_hit", nvf, torch.uint8), ("vf_t", nvf, torch.float64), ("vf_sub", nvf, torch.uint8),
        ("vf_win", nvf, torch.int32), ("ee_hit", nee, torch.uint8), ("ee_t", nee, torch.float64),
        ("ee_sub", nee, torch.uint8), ("ee_win", nee, torch.int32))}
    ctx.check(N.lib().ccd_ctcd_narrowphase_history(
        ctx.ptr, _ptr(off), _ptr(ht), _ptr(hp), V, E, _ptr(vk), None, nvf, _ptr(ek), None, nee, float(eta),
        _ptr(o["vf_hit"]), _ptr(o["vf_t"]), _ptr(o["vf_sub"]), _ptr(o["vf_win"]), _ptr(o["ee_hit"]),
        _ptr(o["ee_t"]), _ptr(o["ee_sub"]), _ptr(o["ee_win"])))
    torch.cuda.synchronize(dev)
    out = {k: v.cpu().numpy() for k, v in o.items()}
    out["vf_hit"] = out["vf_hit"].astype(bool)
    out["ee_hit"] = out["ee_hit"].astype(bool)
    return out


def separating_plane_narrowphase(hist_off, hist_t, hist_pos, vf_keys, ee_keys, eta, vf_eta=None, ee_eta=None,
                                  ctx: Context | None = None):
    """SeparatingPlaneNarrowPhase over stencil lists for a CSR history (SeparatingPlaneNarrowPhase.cpp:9-278).

    Returns dict(vf_status, ee_status, undecided): per-stencil status 1 (collision), 0 (none) or 2
    (undecided: the reference's recursion would not end there)."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    off = to_device(hist_off, torch.int64, ctx)
    ht = to_device(hist_t, torch.float64, ctx)
    hp = to_device(hist_pos, torch.float64, ctx).reshape(-1)
    V, E = off.numel() - 1, ht.numel()
    vk = to_device(vf_keys, torch.int32, ctx).reshape(-1)
    ek = to_device(ee_keys, torch.int32, ctx).reshape(-1)
    nvf, nee = vk.numel() // 4, ek.numel() // 4
    ve = to_device(vf_eta, torch.float64, ctx) if vf_eta is not None else None
    ee = to_device(ee_eta, torch.float64, ctx) if ee_eta is not None else None
    dev = ctx.torch_device
    vs = torch.empty(nvf, dtype=torch.uint8, device=dev)
    es = torch.empty(nee, dtype=torch.uint8, device=dev)
    und = ctypes.c_int64(0)
    ctx.check(N.lib().ccd_separating_plane_narrowphase(
        ctx.ptr, _ptr(off), _ptr(ht), _ptr(hp), V, E, _ptr(vk), _ptr(ve), nvf, _ptr(ek), _ptr(ee), nee, float(eta),
        _ptr(vs), _ptr(es), ctypes.byref(und)))
    return {"vf_status": vs.cpu().numpy(), "ee_status": es.cpu().numpy(), "undecided": und.value}


def _fixed_mask(fixed, V, ctx):
    torch = _torch()
    if fixed is None:
        return None
    m = np.zeros(V, dtype=np.uint8)
    idx = np.asarray(sorted(fixed) if isinstance(fixed, (set, frozenset)) else fixed, dtype=np.int64)
    if idx.size and idx.dtype != bool:
        m[idx] = 1
    return to_device(m, torch.uint8, ctx)


def kdop_candidates(faces, qstart, qend, outer_eta, fixed=None, sort=True, ctx: Context | None = None,
                    hist=None, as_numpy=True):
    """KDOPBroadPhase candidate sets -> (vf, ee) int32 arrays of shape (n, 4).

    hist=(off, t, pos) selects the multi-entry form (qstart/qend are then ignored).  as_numpy=False
    returns device tensors (for sets too large to copy to the host)."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    fd = to_device(faces, torch.int32, ctx).reshape(-1)
    F = fd.numel() // 3
    flags = N.CCD_SORTED if sort else 0
    if hist is None:
        qs = to_device(qstart, torch.float64, ctx).reshape(-1)
        qe = to_device(qend, torch.float64, ctx).reshape(-1)
        V = qs.numel() // 3
    else:
        off = to_device(hist[0], torch.int64, ctx)
        hp = to_device(hist[2], torch.float64, ctx).reshape(-1)
        V = off.numel() - 1
        E = hp.numel() // 3
    fx = _fixed_mask(fixed, V, ctx)
    nv = ctypes.c_int64(0)
    ne = ctypes.c_int64(0)
    cap_v, cap_e = 0, 0
    for _ in range(2):
        vf = torch.empty(max(cap_v, 1) * 4, dtype=torch.int32, device=ctx.torch_device)
        ee = torch.empty(max(cap_e, 1) * 4, dtype=torch.int32, device=ctx.torch_device)
        if hist is None:
            rc = N.lib().ccd_kdop_candidates(ctx.ptr, _ptr(fd), F, _ptr(qs), _ptr(qe), V, float(outer_eta), _ptr(fx),
                                             _ptr(vf), cap_v, ctypes.byref(nv), _ptr(ee), cap_e, ctypes.byref(ne), flags)
        else:
            rc = N.lib().ccd_kdop_candidates_history(ctx.ptr, _ptr(fd), F, _ptr(off), _ptr(hp), V, E,
                                                     float(outer_eta), _ptr(fx), _ptr(vf), cap_v, ctypes.byref(nv),
                                                     _ptr(ee), cap_e, ctypes.byref(ne), flags)
        if rc == N.CCD_E_CAPACITY:
            cap_v, cap_e = nv.value, ne.value
            continue
        ctx.check(rc)
        break
    torch.cuda.synchronize(ctx.torch_device)
    vf, ee = vf[: 4 * nv.value].reshape(-1, 4), ee[: 4 * ne.value].reshape(-1, 4)
    if not as_numpy:
        return vf, ee
    return vf.cpu().numpy(), ee.cpu().numpy()


def detect(faces, qstart, qend, outer_eta, eta, fixed=None, ctx: Context | None = None, cap=None):
    """Broadphase + narrowphase with candidates kept on the GPU.

    Returns dict(vf, vf_toi, ee, ee_toi, vf_candidates, ee_candidates) with colliding stencils."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    fd = to_device(faces, torch.int32, ctx).reshape(-1)
    qs = to_device(qstart, torch.float64, ctx).reshape(-1)
    qe = to_device(qend, torch.float64, ctx).reshape(-1)
    F, V = fd.numel() // 3, qs.numel() // 3
    fx = _fixed_mask(fixed, V, ctx)
    cap_v = cap_e = cap if cap is not None else max(1024, 4 * F)
    c = [ctypes.c_int64(0) for _ in range(4)]
    for _ in range(2):
        dev = ctx.torch_device
        vk = torch.empty(cap_v * 4, dtype=torch.int32, device=dev)
        vt = torch.empty(cap_v, dtype=torch.float64, device=dev)
        ek = torch.empty(cap_e * 4, dtype=torch.int32, device=dev)
        et = torch.empty(cap_e, dtype=torch.float64, device=dev)
        rc = N.lib().ccd_detect(ctx.ptr, _ptr(fd), F, _ptr(qs), _ptr(qe), V, float(outer_eta), float(eta), _ptr(fx),
                                _ptr(vk), _ptr(vt), cap_v, ctypes.byref(c[0]), _ptr(ek), _ptr(et), cap_e,
                                ctypes.byref(c[1]), ctypes.byref(c[2]), ctypes.byref(c[3]))
        if rc == N.CCD_E_CAPACITY:
            cap_v, cap_e = max(c[0].value, 1), max(c[1].value, 1)
            continue
        ctx.check(rc)
        break
    nv, ne = c[0].value, c[1].value
    return {"vf": vk[: 4 * nv].reshape(-1, 4).cpu().numpy(), "vf_toi": vt[:nv].cpu().numpy(),
            "ee": ek[: 4 * ne].reshape(-1, 4).cpu().numpy(), "ee_toi": et[:ne].cpu().numpy(),
            "vf_candidates": c[2].value, "ee_candidates": c[3].value}


# ------------------------------------------------------------------------------------------------
# Reference-interface mirror
# ------------------------------------------------------------------------------------------------
class Mesh:
    """src/Mesh.h:6-12: vertices (3V) and faces (one column per triangle, here one row)."""

    def __init__(self, vertices=None, faces=None):
        self.vertices = np.zeros(0) if vertices is None else np.asarray(vertices, dtype=np.float64).reshape(-1)
        self.faces = np.zeros((0, 3), np.int32) if faces is None else np.asarray(faces, dtype=np.int32).reshape(-1, 3)


class History:
    """src/History.h:24-42: per-vertex (time, position) trajectories, first entry at t=0."""

    def __init__(self, qstart):
        q = np.asarray(qstart, dtype=np.float64).reshape(-1, 3)
        self._entries = [[(0.0, q[i].copy())] for i in range(q.shape[0])]

    def addHistory(self, vert: int, time: float, pos):   # History.cpp:21-26
        self._entries[vert].append((float(time), np.asarray(pos, dtype=np.float64).reshape(3).copy()))

    def finishHistory(self, qend):   # History.cpp:27-39
        q = np.asarray(qend, dtype=np.float64).reshape(-1, 3)
        for i in range(q.shape[0]):
            self._entries[i].append((1.0, q[i].copy()))

    def countHistoryEntries(self) -> int:
        return sum(len(e) for e in self._entries)

    def is_two_entry(self) -> bool:
        return all(len(e) == 2 and e[0][0] == 0.0 and e[1][0] == 1.0 for e in self._entries)

    def endpoints(self):
        qs = np.stack([e[0][1] for e in self._entries])
        qe = np.stack([e[-1][1] for e in self._entries])
        return qs, qe

    def csr(self):
        off = np.zeros(len(self._entries) + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(e) for e in self._entries])
        ht = np.array([t for e in self._entries for (t, _) in e], dtype=np.float64)
        hp = np.array([p for e in self._entries for (_, p) in e], dtype=np.float64).reshape(-1)
        return off, ht, hp


class VertexFaceStencil(tuple):
    """Stencils.h:7-78: (p, q0, q1, q2) with the face's vertices ascending."""

    def __new__(cls, p, q0, q1, q2):
        a = sorted((int(q0), int(q1), int(q2)))
        return super().__new__(cls, (int(p), a[0], a[1], a[2]))


class EdgeEdgeStencil(tuple):
    """Stencils.h:80-138: each edge ascending, the edge with the smaller first vertex first."""

    def __new__(cls, p0, p1, q0, q1):
        p = sorted((int(p0), int(p1)))
        q = sorted((int(q0), int(q1)))
        if p[0] > q[0]:
            p, q = q, p
        return super().__new__(cls, (p[0], p[1], q[0], q[1]))


class KDOPBroadPhase:
    """src/KDOPBroadPhase.h:49-61 -- BroadPhase plugin (RetrospectiveDetection.h:10-15)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx

    def findCollisionCandidates(self, h: History, m: Mesh, outerEta: float, vfs: set, ees: set, fixedVerts=()):
        vfs.clear()   # KDOPBroadPhase.cpp:36-37
        ees.clear()
        if h.is_two_entry():
            qs, qe = h.endpoints()
            vf, ee = kdop_candidates(m.faces, qs, qe, outerEta, fixed=set(fixedVerts) or None, ctx=self.ctx)
        else:
            vf, ee = kdop_candidates(m.faces, None, None, outerEta, fixed=set(fixedVerts) or None, ctx=self.ctx,
                                     hist=h.csr())
        vfs.update(tuple(int(x) for x in r) for r in vf)
        ees.update(tuple(int(x) for x in r) for r in ee)


class CTCDNarrowPhase:
    """src/CTCDNarrowPhase.h:6-15 -- NarrowPhase plugin (RetrospectiveDetection.h:17-23).

    candidateVFS / candidateEES are sets of (stencil, eta) pairs; colliding stencils are inserted into
    vfs / ees (which are not cleared, CTCDNarrowPhase.cpp:9-22)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx

    def findCollisions(self, h: History, candidateVFS, candidateEES, vfs: set, ees: set):
        cv = sorted(candidateVFS)
        ce = sorted(candidateEES)
        vk = np.array([s for s, _ in cv], dtype=np.int32).reshape(-1, 4)
        ve = np.array([e for _, e in cv], dtype=np.float64)
        ek = np.array([s for s, _ in ce], dtype=np.int32).reshape(-1, 4)
        eet = np.array([e for _, e in ce], dtype=np.float64)
        if h.is_two_entry():
            qs, qe = h.endpoints()
            r = narrowphase(qs, qe, vk, ek, 0.0, vf_eta=ve if len(ve) else None, ee_eta=eet if len(eet) else None,
                            ctx=self.ctx)
        else:
            off, ht, hp = h.csr()
            if len(set(ve.tolist()) | set(eet.tolist())) > 1:
                raise NotImplementedError("per-stencil eta with multi-entry histories")
            eta = float(ve[0]) if len(ve) else (float(eet[0]) if len(eet) else 0.0)
            r = narrowphase_history(off, ht, hp, vk, ek, eta, ctx=self.ctx)
        for (s, _), hit in zip(cv, r["vf_hit"]):
            if hit:
                vfs.add(tuple(s))
        for (s, _), hit in zip(ce, r["ee_hit"]):
            if hit:
                ees.add(tuple(s))


class SeparatingPlaneNarrowPhase:
    """SeparatingPlaneNarrowPhase (src/SeparatingPlaneNarrowPhase.h:7-22): findCollisions with the same
    arguments as CTCDNarrowPhase's; colliding stencils are inserted into vfs / ees (not cleared).  A
    stencil on which the reference's recursion would not end raises RuntimeError (the reference
    overflows its stack there)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx

    def findCollisions(self, h: History, candidateVFS, candidateEES, vfs: set, ees: set):
        cv = sorted(candidateVFS)
        ce = sorted(candidateEES)
        vk = np.array([s for s, _ in cv], dtype=np.int32).reshape(-1, 4)
        ve = np.array([e for _, e in cv], dtype=np.float64)
        ek = np.array([s for s, _ in ce], dtype=np.int32).reshape(-1, 4)
        eet = np.array([e for _, e in ce], dtype=np.float64)
        off, ht, hp = h.csr()
        r = separating_plane_narrowphase(off, ht, hp, vk, ek, 0.0, vf_eta=ve if len(ve) else None,
                                         ee_eta=eet if len(eet) else None, ctx=self.ctx)
        if r["undecided"]:
            bad = [tuple(s) for (s, _), st in zip(cv, r["vf_status"]) if st == 2] + \
                  [tuple(s) for (s, _), st in zip(ce, r["ee_status"]) if st == 2]
            raise RuntimeError(f"SeparatingPlaneNarrowPhase: the recursion does not terminate for {len(bad)} "
                               f"stencil(s), e.g. {bad[:3]}")
        for (s, _), st in zip(cv, r["vf_status"]):
            if st == 1:
                vfs.add(tuple(s))
        for (s, _), st in zip(ce, r["ee_status"]):
            if st == 1:
                ees.add(tuple(s))


class CTCD:
    """include/CTCD.h:28-79 -- the primitive statics; each returns (collides, t) with t None when false."""

    @staticmethod
    def _one(kind, pts, eta):
        q = np.concatenate([np.asarray(p, dtype=np.float64).reshape(3) for p in pts])[None, :]
        hit, t = primitives(kind, q, eta)
        return bool(hit[0]), (float(t[0]) if hit[0] else None)

    @staticmethod
    def edgeEdgeCTCD(q0start, p0start, q1start, p1start, q0end, p0end, q1end, p1end, eta):
        return CTCD._one(N.PRIM_EDGE_EDGE, (q0start, p0start, q1start, p1start, q0end, p0end, q1end, p1end), eta)

    @staticmethod
    def vertexFaceCTCD(q0start, q1start, q2start, q3start, q0end, q1end, q2end, q3end, eta):
        return CTCD._one(N.PRIM_VERTEX_FACE, (q0start, q1start, q2start, q3start, q0end, q1end, q2end, q3end), eta)

    @staticmethod
    def vertexEdgeCTCD(q0start, q1start, q2start, q0end, q1end, q2end, eta):
        return CTCD._one(N.PRIM_VERTEX_EDGE, (q0start, q1start, q2start, q0end, q1end, q2end), eta)

    @staticmethod
    def vertexVertexCTCD(q1start, q2start, q1end, q2end, eta):
        return CTCD._one(N.PRIM_VERTEX_VERTEX, (q1start, q2start, q1end, q2end), eta)


# ------------------------------------------------------------------------------------------------
# Proximity
# ------------------------------------------------------------------------------------------------
def distance_primitives(kind: int, q, eta: float = 0.0, ctx: Context | None = None):
    """The Distance statics on n items (q: (n, 12)); returns (flag, vec (n,3), bary (n,4)) as numpy."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    qd = to_device(q, torch.float64, ctx).reshape(-1)
    n = qd.numel() // 12
    dev = ctx.torch_device
    flag = torch.zeros(n, dtype=torch.uint8, device=dev)
    vec = torch.zeros(3 * n, dtype=torch.float64, device=dev)
    bary = torch.zeros(4 * n, dtype=torch.float64, device=dev)
    ctx.check(N.lib().ccd_distance_primitives(ctx.ptr, int(kind), _ptr(qd), float(eta), n, _ptr(flag), _ptr(vec),
                                              _ptr(bary)))
    torch.cuda.synchronize(dev)
    return flag.cpu().numpy().astype(bool), vec.cpu().numpy().reshape(-1, 3), bary.cpu().numpy().reshape(-1, 4)


def mesh_self_distance(verts, faces, fixed=None, ctx: Context | None = None):
    """Distance::meshSelfDistance -> (distance, n_vf_candidates, n_ee_candidates)."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    v = to_device(verts, torch.float64, ctx).reshape(-1)
    f = to_device(faces, torch.int32, ctx).reshape(-1)
    V, F = v.numel() // 3, f.numel() // 3
    fx = _fixed_mask(fixed, V, ctx)
    d = ctypes.c_double(0.0)
    nv = ctypes.c_int64(0)
    ne = ctypes.c_int64(0)
    ctx.check(N.lib().ccd_mesh_self_distance(ctx.ptr, _ptr(v), V, _ptr(f), F, _ptr(fx), ctypes.byref(d),
                                             ctypes.byref(nv), ctypes.byref(ne)))
    return d.value, nv.value, ne.value


# ------------------------------------------------------------------------------------------------
# Velocity filter
# ------------------------------------------------------------------------------------------------
def velocity_filter(qstart, qend, faces, inv_masses, outer_eta, inner_eta, base_stiffness=1e-4, base_substep=1e-4,
                    cor=0.1, max_rollbacks=1000, verbose=False, ctx: Context | None = None):
    """VelocityFilter::velocityFilter on host arrays -> (result, qend_out, stats).

    result is the reference's return value (outer iterations taken, or -1 / -2 / -3); qend_out is a new
    float64 array, the filtered end positions when result >= 0 and a copy of qend otherwise; stats is a
    dict of the ccd_vf_stats counters."""
    ctx = ctx or default_context()
    ctx.bind_stream()
    qs = np.ascontiguousarray(qstart, dtype=np.float64).reshape(-1)
    qe = np.array(qend, dtype=np.float64).reshape(-1)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
    im = np.ascontiguousarray(inv_masses, dtype=np.float64).reshape(-1)
    V, F = len(qs) // 3, len(f) // 3
    if len(qs) != 3 * V or len(qe) != 3 * V or len(im) != 3 * V or len(f) != 3 * F:
        raise ValueError("qstart, qend and inv_masses need 3*V entries and faces 3*F")
    res = ctypes.c_int(0)
    st = N.VFStats()
    dp = ctypes.POINTER(ctypes.c_double)
    ctx.check(N.lib().ccd_velocity_filter(ctx.ptr, qs.ctypes.data_as(dp), qe.ctypes.data_as(dp), V,
                                          f.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), F, im.ctypes.data_as(dp),
                                          float(outer_eta), float(inner_eta), float(base_stiffness),
                                          float(base_substep), float(cor), int(max_rollbacks), int(bool(verbose)),
                                          ctypes.byref(res), ctypes.byref(st)))
    return res.value, qe, {name: getattr(st, name) for name, _ in N.VFStats._fields_}


class VelocityFilter:
    """VelocityFilter (include/VelocityFilter.h:7-45): the return codes and the static velocityFilter."""

    DANGLING_VERTICES = -1
    STARTING_STATE_INTERSECTS = -2
    MAXROLLBACKS_EXCEEDED = -3

    @staticmethod
    def velocityFilter(qstart, qend, faces, masses, outerEta, innerEta, baseStiffness=1e-4, baseSubstepSize=1e-4,
                       CoR=0.1, maxRollbacks=1000):
        """As the reference: qend (a float64 numpy array) is overwritten with the filtered end positions on
        success; returns the number of outer iterations or a negative code.  masses are inverse masses
        per DOF.  Prints the reference's progress lines."""
        r, qe, _ = velocity_filter(qstart, qend, faces, masses, outerEta, innerEta, baseStiffness, baseSubstepSize,
                                   CoR, maxRollbacks, verbose=True)
        if r >= 0:
            np.copyto(np.asarray(qend).reshape(-1), qe)
        return r


# ------------------------------------------------------------------------------------------------
# Batched instances
# ------------------------------------------------------------------------------------------------
class Batch:
    """Independent meshes packed for ccd_batch_detect: faces with global vertex indices, face offsets
    per instance and vertex offsets (to map results back)."""

    def __init__(self, scenes_list, ctx: Context | None = None):
        torch = _torch()
        self.ctx = ctx or default_context()
        fs, qs, qe, foff, voff = [], [], [], [0], [0]
        for s in scenes_list:
            fs.append(np.asarray(s["faces"], dtype=np.int32) + voff[-1])
            qs.append(np.asarray(s["qstart"], dtype=np.float64))
            qe.append(np.asarray(s["qend"], dtype=np.float64))
            foff.append(foff[-1] + len(s["faces"]))
            voff.append(voff[-1] + len(s["qstart"]))
        self.ninst = len(scenes_list)
        self.face_off = np.asarray(foff, dtype=np.int64)
        self.vert_off = np.asarray(voff, dtype=np.int64)
        self.F, self.V = int(foff[-1]), int(voff[-1])
        c = self.ctx
        self.faces = to_device(np.concatenate(fs) if fs else np.zeros((0, 3), np.int32), torch.int32, c).reshape(-1)
        self.qs = to_device(np.concatenate(qs) if qs else np.zeros((0, 3)), torch.float64, c).reshape(-1)
        self.qe = to_device(np.concatenate(qe) if qe else np.zeros((0, 3)), torch.float64, c).reshape(-1)
        self.off_d = to_device(self.face_off, torch.int64, c)
        self.cap = max(1024, 2 * self.F)
        self._alloc()

    def _alloc(self):
        torch = _torch()
        dev = self.ctx.torch_device
        self.vk = torch.empty(4 * self.cap, dtype=torch.int32, device=dev)
        self.vt = torch.empty(self.cap, dtype=torch.float64, device=dev)
        self.ek = torch.empty(4 * self.cap, dtype=torch.int32, device=dev)
        self.et = torch.empty(self.cap, dtype=torch.float64, device=dev)

    def instance_of_vertex(self, v):
        return np.searchsorted(self.vert_off, v, side="right") - 1

    def tune(self, outer_eta, eta, cache=None):
        """Tune the persistent narrowphase (queue shape and stencil grouping per kind) on this batch's
        candidates and install the result on the batch's context.  With `cache` (a JSON file path) the
        shape stored there for this workload shape (GPU architecture and compute units, faces, vertices,
        instances) is installed
        instead, and a newly tuned shape is stored.  Returns (shape, 'cache' or 'tuned')."""
        import json
        import os

        torch = _torch()
        props = torch.cuda.get_device_properties(self.ctx.torch_device)
        arch = (getattr(props, "gcnArchName", "") or props.name).split(":")[0]
        key = f"{arch}|cus={props.multi_processor_count}|faces={self.F}|verts={self.V}|instances={self.ninst}"
        table = {}
        if cache and os.path.exists(cache):
            with open(cache) as fh:
                table = json.load(fh)
        if key in table:
            set_queue_shape(table[key], self.ctx)
            return list(table[key]), "cache"
        self.detect(outer_eta, eta, graph=False)   # the candidates the tuner times
        self.ctx.bind_stream()
        shape = (ctypes.c_int * 8)()
        self.ctx.check(N.lib().ccd_autotune(self.ctx.ptr, _ptr(self.qs), _ptr(self.qe), float(eta), shape))
        shape = list(shape)
        if cache:
            table[key] = shape
            tmp = cache + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(table, fh, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return shape, "tuned"

    def detect(self, outer_eta, eta, graph=False):
        """One ccd_batch_detect call -> dict(vf, vf_toi, ee, ee_toi, vf_candidates, ee_candidates) with global
        vertex indices (use instance_of_vertex to split)."""
        ctx = self.ctx
        ctx.bind_stream()
        c = [ctypes.c_int64(0) for _ in range(4)]
        for _ in range(2):
            rc = N.lib().ccd_batch_detect(ctx.ptr, self.ninst, _ptr(self.off_d), _ptr(self.faces), self.F, _ptr(self.qs),
                                          _ptr(self.qe), self.V, float(outer_eta), float(eta), _ptr(self.vk),
                                          _ptr(self.vt), self.cap, ctypes.byref(c[0]), _ptr(self.ek), _ptr(self.et),
                                          self.cap, ctypes.byref(c[1]), ctypes.byref(c[2]), ctypes.byref(c[3]),
                                          N.CCD_GRAPH if graph else 0)
            if rc == N.CCD_E_CAPACITY:
                self.cap = max(c[0].value, c[1].value) + 1024
                self._alloc()
                continue
            ctx.check(rc)
            break
        nv, ne = c[0].value, c[1].value
        return {"vf": self.vk[: 4 * nv].reshape(-1, 4).cpu().numpy(), "vf_toi": self.vt[:nv].cpu().numpy(),
                "ee": self.ek[: 4 * ne].reshape(-1, 4).cpu().numpy(), "ee_toi": self.et[:ne].cpu().numpy(),
                "vf_candidates": c[2].value, "ee_candidates": c[3].value}


def autotune(qstart, qend, eta, ctx: Context | None = None):
    """Tune the persistent narrowphase on the last detection's candidates; returns the 8-int shape
    (block, tile, waves/CU per kind, then the grouping flag per kind)."""
    torch = _torch()
    ctx = ctx or default_context()
    ctx.bind_stream()
    qs = to_device(qstart, torch.float64, ctx).reshape(-1)
    qe = to_device(qend, torch.float64, ctx).reshape(-1)
    shape = (ctypes.c_int * 8)()
    ctx.check(N.lib().ccd_autotune(ctx.ptr, _ptr(qs), _ptr(qe), float(eta), shape))
    return list(shape)


def set_queue_shape(shape, ctx: Context | None = None):
    """Install an 8-int shape from autotune (e.g. a cached one)."""
    ctx = ctx or default_context()
    arr = (ctypes.c_int * 8)(*[int(x) for x in shape])
    ctx.check(N.lib().ccd_set_queue_shape(ctx.ptr, arr))


def set_edge_edge_split(on: bool, ctx: Context | None = None):
    """Two-pass (True, the default) or one-pass edge-edge narrowphase; the results are the same."""
    ctx = ctx or default_context()
    ctx.check(N.lib().ccd_set_edge_edge_split(ctx.ptr, 1 if on else 0))


# ------------------------------------------------------------------------------------------------
# Serving
# ------------------------------------------------------------------------------------------------
class Server:
    """Device-resident sessions behind a continuous-batching request loop (ccd_server_* in
    include/ccd_amd.h).  Sessions refit their tree per request; requests of different sessions
    overlap on the device."""

    def __init__(self, hbm_budget_bytes=8 << 30, max_inflight=8, device=0):
        _torch()
        self._p = ctypes.c_void_p()
        rc = N.lib().ccd_server_create(device, int(hbm_budget_bytes), int(max_inflight), ctypes.byref(self._p))
        if rc != N.CCD_OK:
            raise N.CCDError(rc, "ccd_server_create failed")
        self.nverts = {}
        self._vf_hits = {}

    def close(self):
        if self._p is not None and self._p.value:
            N.lib().ccd_server_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def open_session(self, faces, qstart, qend, outer_eta, eta) -> int:
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        a = np.ascontiguousarray(qstart, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(qend, dtype=np.float64).reshape(-1)
        sid = ctypes.c_int(-1)
        rc = N.lib().ccd_server_open_session(self._p, f.ctypes.data_as(ctypes.c_void_p), len(f) // 3,
                                             a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                                             len(a) // 3, float(outer_eta), float(eta), ctypes.byref(sid))
        if rc != N.CCD_OK:
            raise N.CCDError(rc, "ccd_server_open_session failed")
        self.nverts[sid.value] = len(a) // 3
        return sid.value

    def submit(self, session: int, qend_next, request_id: int):
        q = np.ascontiguousarray(qend_next, dtype=np.float64).reshape(-1)
        assert len(q) == 3 * self.nverts[session]
        rc = N.lib().ccd_server_submit(self._p, int(session), q.ctypes.data_as(ctypes.c_void_p), int(request_id))
        if rc != N.CCD_OK:
            raise N.CCDError(rc, "ccd_server_submit failed")

    def poll(self, max_results=4096):
        buf = (N.RequestResult * max_results)()
        n = ctypes.c_int(0)
        N.lib().ccd_server_poll(self._p, buf, max_results, ctypes.byref(n))
        out = [{f: getattr(buf[i], f) for f, _ in N.RequestResult._fields_} for i in range(n.value)]
        for r in out:
            self._vf_hits[r["request_id"]] = r["vf_hits"]
        return out

    def drain(self):
        N.lib().ccd_server_drain(self._p)

    def stats(self):
        s = (ctypes.c_int64 * 10)()
        N.lib().ccd_server_stats(self._p, s)
        return dict(zip(("submitted", "completed", "pending", "inflight", "pool_bytes", "resident_bytes",
                         "reruns", "sessions", "mem_waits", "max_request_bytes"), list(s)))

    def keep_hits(self, on=True):
        N.lib().ccd_server_keep_hits(self._p, int(on))

    def fetch_hits(self, request_id):
        """-> (vf keys (n,4), vf toi, ee keys (m,4), ee toi) of a request that completed while keep_hits
        was on and has been returned by poll()."""
        n = ctypes.c_int64(0)
        N.lib().ccd_server_fetch_hits(self._p, int(request_id), None, 0, ctypes.byref(n))
        buf = np.zeros(n.value, dtype=np.int32)
        rc = N.lib().ccd_server_fetch_hits(self._p, int(request_id), buf.ctypes.data_as(ctypes.c_void_p), n.value,
                                           ctypes.byref(n))
        if rc != N.CCD_OK:
            raise N.CCDError(rc, "ccd_server_fetch_hits failed")
        m = n.value // 6   # per stencil: 4 key ints, then (after all keys) 2 ints of FP64 TOI
        nv = self._vf_hits[request_id]
        keys = buf[: 4 * m].reshape(-1, 4)
        toi = buf[4 * m:].view(np.float64)
        return keys[:nv], toi[:nv], keys[nv:], toi[nv:]
