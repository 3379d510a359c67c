"""Seeded synthetic CCD workloads (swept triangle meshes in FP64).

Every scene is a dict with
    faces  : int32 (F, 3)   -- one row per triangle (a column of the reference's Eigen::Matrix3Xi)
    qstart : float64 (V, 3) -- positions at t = 0 (the reference's History(qstart), History.cpp:8-19)
    qend   : float64 (V, 3) -- positions at t = 1 (History::finishHistory, History.cpp:27-39)
    eta    : float          -- used as both the broadphase outerEta and the per-stencil eta
    name   : str
Random draws use numpy.random.default_rng(seed) (PCG64), so a scene is fixed by its parameters.
"""
from __future__ import annotations

import functools
import hashlib
import math

import numpy as np

# ------------------------------------------------------------------------------------------------
# Mesh builders
# ------------------------------------------------------------------------------------------------


def icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


def icosphere(levels: int, radius: float = 1.0):
    """Icosahedron with `levels` midpoint subdivisions, projected to the sphere: 10*4^k+2 V, 20*4^k F."""
    v, f = _unit_icosphere(levels)
    return v * radius, f.copy()


@functools.lru_cache(maxsize=None)
def _unit_icosphere_cached(levels: int):
    return _build_icosphere(levels)


def _unit_icosphere(levels: int):
    v, f = _unit_icosphere_cached(levels)
    return v.copy(), f


def _build_icosphere(levels: int):
    v, f = icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(levels):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key in cache:
                return cache[key]
            p = (np.asarray(verts[a]) + np.asarray(verts[b])) * 0.5
            p /= np.linalg.norm(p)
            verts.append(tuple(p))
            cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = np.asarray(nf, dtype=np.int32)
    return np.asarray(verts, dtype=np.float64), f


def grid(n: int, x0=0.0, x1=1.0, y0=0.0, y1=1.0):
    """n x n vertex grid; vertex id = n*j + i; quad (i,j) -> triangles (a,b,d), (a,d,c)."""
    xs = x0 + (x1 - x0) * np.arange(n) / (n - 1)
    ys = y0 + (y1 - y0) * np.arange(n) / (n - 1)
    X, Y = np.meshgrid(xs, ys)   # X[j, i] = xs[i]
    v = np.stack([X.reshape(-1), Y.reshape(-1), np.zeros(n * n)], axis=1)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1))
    i = i.reshape(-1)
    j = j.reshape(-1)
    a = n * j + i
    b = a + 1
    c = a + n
    d = c + 1
    f = np.empty((2 * len(a), 3), dtype=np.int32)
    f[0::2] = np.stack([a, b, d], axis=1)
    f[1::2] = np.stack([a, d, c], axis=1)
    return v, f


def _scene(name, faces, qstart, qend, eta):
    return {"name": name, "faces": np.ascontiguousarray(faces, dtype=np.int32),
            "qstart": np.ascontiguousarray(qstart, dtype=np.float64),
            "qend": np.ascontiguousarray(qend, dtype=np.float64), "eta": float(eta)}


# ------------------------------------------------------------------------------------------------
# Workloads
# ------------------------------------------------------------------------------------------------


def two_spheres(seed: int = 0, levels: int = 3, offset: float = 2.05, shift: float = 0.3,
                jitter: float = 0.01, eta: float = 1e-5):
    """Two radius-1 icospheres (642 V / 1,280 F each at levels=3) in one mesh; B starts at +offset
    along x.  End = start + rigid -shift (A) / +shift (B) along x + N(0, jitter^2) per coordinate,
    drawn as one (V, 3) array in vertex order."""
    v, f = icosphere(levels)
    nv = len(v)
    qs = np.concatenate([v, v + np.array([offset, 0.0, 0.0])])
    faces = np.concatenate([f, f + nv])
    rng = np.random.default_rng(seed)
    noise = rng.normal(0.0, jitter, size=qs.shape)
    disp = np.zeros_like(qs)
    disp[:nv, 0] = -shift
    disp[nv:, 0] = shift
    return _scene(f"two_spheres_l{levels}_s{seed}", faces, qs, qs + disp + noise, eta)


def converging_spheres(seed: int = 10, levels: int = 2, eta: float = 1e-5):
    """A colliding variant of two_spheres: B starts at +2.05 and the spheres approach by 0.3 each."""
    return two_spheres(seed=seed, levels=levels, offset=2.05, shift=-0.3, jitter=0.01, eta=eta)


def cloth(n: int = 256, seed: int = 1, drop: float = 0.02, noise: float = 0.004, eta: float = 1e-6):
    """Draped-fold cloth self-CCD: n x n grid on [0,1]^2, z = 0.15 sin(6x) cos(5y); end = start +
    (0, 0, -drop) + N(0, noise^2) per coordinate.  n=256: 65,536 V / 130,050 F."""
    v, f = grid(n)
    v[:, 2] = 0.15 * np.sin(6.0 * v[:, 0]) * np.cos(5.0 * v[:, 1])
    rng = np.random.default_rng(seed)
    qe = v + np.array([0.0, 0.0, -drop]) + rng.normal(0.0, noise, size=v.shape)
    return _scene(f"cloth_{n}_s{seed}", f, v, qe, eta)


def crumpled_sheet(n: int = 1024, seed: int = 3, eta: float = 1e-7):
    """Large self-CCD: n x n sheet, start z = sum of 8 folds a_k sin(f_k (cos th_k x + sin th_k y) + ph_k)
    with a_k ~ U(0.01, 0.1), f_k ~ U(2, 20); end = start + per-component sum of 8 sinusoids of
    amplitude 0.01 + N(0, 0.001^2).  n=1024: 1,048,576 V / 2,093,058 F."""
    v, f = grid(n)
    rng = np.random.default_rng(seed)
    x, y = v[:, 0], v[:, 1]
    z = np.zeros(len(v))
    for _ in range(8):
        a, fr = rng.uniform(0.01, 0.1), rng.uniform(2, 20)
        th, ph = rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)
        z += a * np.sin(fr * (np.cos(th) * x + np.sin(th) * y) + ph)
    v[:, 2] = z
    disp = np.zeros_like(v)
    for c in range(3):
        for _ in range(8):
            fr = rng.uniform(2, 20)
            th, ph = rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)
            disp[:, c] += 0.01 * np.sin(fr * (np.cos(th) * x + np.sin(th) * y) + ph)
    qe = v + disp + rng.normal(0.0, 0.001, size=v.shape)
    return _scene(f"crumpled_{n}_s{seed}", f, v, qe, eta)


def falling_sheets(n: int = 4, seed: int = 0, gap: float = 0.1, drop: float = 0.2, drift: float = 0.05,
                   jitter: float = 0.02, outer_eta: float = 0.05, inner_eta: float = 0.01):
    """A velocity-filter contact problem: two jittered n x n grid sheets, the upper one (gap above,
    shifted sideways) falling by `drop` through the lower one while drifting by `drift` in x.  The
    start state is separated (closest pair ~gap apart), the end state interpenetrates.  Adds the
    filter's layer thicknesses outer_eta / inner_eta to the scene dict (eta = outer_eta)."""
    rng = np.random.default_rng(seed)
    sheets = []
    for z, off in ((0.0, (0.0, 0.0)), (gap, (0.13, 0.07))):
        v, f = grid(n)
        v[:, 0] += off[0] + rng.uniform(-jitter, jitter, n * n)
        v[:, 1] += off[1] + rng.uniform(-jitter, jitter, n * n)
        v[:, 2] = z + 0.1 * rng.uniform(-jitter, jitter, n * n)
        sheets.append((v, f))
    (va, fa), (vb, fb) = sheets
    qstart = np.concatenate([va, vb])
    qend = qstart.copy()
    qend[len(va):, 2] -= drop
    qend[len(va):, 0] += drift
    s = _scene(f"falling_sheets{n}_s{seed}", np.concatenate([fa, fb + len(va)]), qstart, qend, outer_eta)
    s["outer_eta"] = float(outer_eta)
    s["inner_eta"] = float(inner_eta)
    return s


def vertex_bounce(height: float = 0.1, drop: float = 0.2, outer_eta: float = 0.01, inner_eta: float = 0.005):
    """A single contact: a large triangle lying in z = 0 (vertices 0-2) and a small upright triangle whose
    lowest vertex (3) starts `height` above it and falls by `drop` with the rest of its triangle, whose
    other vertices stay far above.  Only the vertex-face stencil (3; 0, 1, 2) comes within outer_eta."""
    qstart = np.array([[-10.0, -10.0, 0.0], [10.0, -10.0, 0.0], [0.0, 10.0, 0.0],
                       [0.1, 0.2, height], [0.1, 0.5, height + 0.5], [0.4, 0.2, height + 0.5]])
    qend = qstart.copy()
    qend[3:, 2] -= drop
    s = _scene("vertex_bounce", np.array([[0, 1, 2], [3, 4, 5]]), qstart, qend, outer_eta)
    s["outer_eta"] = float(outer_eta)
    s["inner_eta"] = float(inner_eta)
    return s


def _rotation(rng, max_angle):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = rng.uniform(0.0, max_angle)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)


def _instance_mesh(rng, kind, target):
    if kind == 0:   # icosphere snapped to 20*4^k triangles, k = 1..5
        k = int(np.clip(round(math.log(target / 20.0, 4)), 1, 5))
        return icosphere(k)
    n = max(3, int(round(math.sqrt(target / 2.0))) + 1)   # noisy grid, 2(n-1)^2 ~ target
    v, f = grid(n, -1.0, 1.0, -1.0, 1.0)
    v[:, 2] = rng.normal(0.0, 0.005, size=n * n)
    return v, f


def batched_instances(count: int = 512, seed: int = 2, eta: float = 1e-6, tri_lo: float = 80,
                      tri_hi: float = 20480):
    """Independent two-mesh CCD instances of uneven size: each instance is a pair of icospheres or
    noisy grids (50/50) with log-uniform triangle targets in [tri_lo, tri_hi] per mesh; mesh B starts
    at a centre separation U(1.9, 2.2) radii along a random direction; each mesh moves by a random
    rigid motion (axis uniform, angle U(0, 0.3), translation U(-0.3, 0.3)^3) plus N(0, 0.005^2)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        meshes = []
        for m in range(2):
            kind = int(rng.integers(0, 2))
            target = math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi)))
            meshes.append(_instance_mesh(rng, kind, target))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        sep = rng.uniform(1.9, 2.2)
        vs, fs, qe = [], [], []
        base = 0
        for m, (v, f) in enumerate(meshes):
            p = v + (sep * d if m == 1 else 0.0)
            c = p.mean(axis=0)
            R = _rotation(rng, 0.3)
            tr = rng.uniform(-0.3, 0.3, size=3)
            e = (p - c) @ R.T + c + tr + rng.normal(0.0, 0.005, size=p.shape)
            vs.append(p)
            qe.append(e)
            fs.append(f + base)
            base += len(v)
        out.append(_scene(f"inst{k}_s{seed}", np.concatenate(fs), np.concatenate(vs), np.concatenate(qe), eta))
    return out


def digest(scene) -> str:
    """SHA-256 over the scene's blobs (faces int32, qstart, qend float64, little-endian) and eta."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(scene["faces"], dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(scene["qstart"], dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(scene["qend"], dtype="<f8").tobytes())
    h.update(np.float64(scene["eta"]).astype("<f8").tobytes())
    return h.hexdigest()


def served_sessions(count: int = 256, seed: int = 4, tri_lo: float = 1000, tri_hi: float = 500000,
                    spacing: float = 0.02, eta: float = 1e-6):
    """Resident sessions for the served workload: `count` sessions with log-uniform triangle counts in
    [tri_lo, tri_hi]; each is (50/50) a self-colliding draped sheet (as `cloth`) or a pair of noisy
    sheets 2*spacing apart.  Grid spacing is fixed at `spacing` so that density does not grow with
    size.  Start = end = the rest shape; requests move the end (see `served_requests`)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        target = math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi)))
        kind = int(rng.integers(0, 2))
        if kind == 0:
            n = max(3, int(round(math.sqrt(target / 2.0))) + 1)
            L = spacing * (n - 1)
            v, f = grid(n, 0.0, L, 0.0, L)
            v[:, 2] = 0.15 * L * np.sin(6.0 * v[:, 0] / L) * np.cos(5.0 * v[:, 1] / L)
        else:
            n = max(3, int(round(math.sqrt(target / 4.0))) + 1)
            L = spacing * (n - 1)
            a, fa = grid(n, 0.0, L, 0.0, L)
            b = a.copy()
            a[:, 2] = rng.normal(0.0, 0.1 * spacing, size=len(a))
            b[:, 2] = 2.0 * spacing + rng.normal(0.0, 0.1 * spacing, size=len(b))
            v = np.concatenate([a, b])
            f = np.concatenate([fa, fa + len(a)])
        out.append(_scene(f"session{k}_s{seed}", f, v, v.copy(), eta))
    return out


class NoiseBank:
    """N(0, sigma^2) draws for request generation at serving rates: one bank of 2^24 normals (seed),
    read at a random offset per request (a fresh draw per request would cost the client more than
    the server spends on it)."""

    def __init__(self, sigma: float = 0.002, seed: int = 4, size: int = 1 << 24):
        self.rng = np.random.default_rng(seed)
        self.bank = self.rng.normal(0.0, sigma, size=size)

    def draw(self, n: int):
        if n > len(self.bank):
            return self.rng.normal(0.0, self.bank.std(), size=n)
        o = int(self.rng.integers(0, len(self.bank) - n + 1))
        return self.bank[o:o + n]
