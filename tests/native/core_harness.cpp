// Test-only host build of the product's FP64 core (collisiondetection_amd/csrc/ccd_core.h).  It lets
// the CPU test suite compare the core's algorithm with the oracle without a GPU, and the GPU tests
// separate "the algorithm differs from the reference" from "the device arithmetic differs from the
// host".  Built by tests/native/build.sh into tests/_build/ (OpenMP over stencils, so that whole
// candidate sets of the large scenes can be checked); never shipped.
#include <cstdint>

#include "../../collisiondetection_amd/csrc/ccd_core.h"

using ccd::V3;

static V3 ld(const double *p) { return ccd::mk(p[0], p[1], p[2]); }

// Separating-plane leaves: 1 = pruned cascades (ccd::NoCount), 0 = the full cascade (ccd::FullCascade, what
// the device kernels run).  Both give the same statuses by construction.
static int g_sp_prune = 1;
extern "C" void core_set_sp_pruning(int on) { g_sp_prune = on; }

template <class C>
static bool primitive(int kind, const double *q, double eta, double &tt, C c)
{
    if (kind == 0)
        return ccd::vertex_face(ld(q), ld(q + 3), ld(q + 6), ld(q + 9), ld(q + 12), ld(q + 15), ld(q + 18), ld(q + 21), eta, tt, c);
    if (kind == 1)
        return ccd::edge_edge(ld(q), ld(q + 3), ld(q + 6), ld(q + 9), ld(q + 12), ld(q + 15), ld(q + 18), ld(q + 21), eta, tt, c);
    if (kind == 2) return ccd::vertex_edge(ld(q), ld(q + 3), ld(q + 6), ld(q + 9), ld(q + 12), ld(q + 15), eta, tt, c);
    return ccd::vertex_vertex(ld(q), ld(q + 3), ld(q + 6), ld(q + 9), eta, tt, c);
}

extern "C" int core_primitive(int kind, const double *q, double eta, double *t)
{
    double tt = 0.0;
    bool r = primitive(kind, q, eta, tt, ccd::NoCount());
    *t = r ? tt : 0.0;
    return r ? 1 : 0;
}

extern "C" void core_narrowphase(const double *qs, const double *qe, const int32_t *keys, int64_t n, int is_ee,
                                 double eta, int32_t *hit, double *t, int32_t *sub)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        V3 s[4], e[4];
        for (int j = 0; j < 4; j++) {
            const double *a = qs + 3 * (int64_t)keys[4 * i + j];
            const double *b = qe + 3 * (int64_t)keys[4 * i + j];
            e[j] = ld(b);
            s[j] = ccd::mk(ccd::stitch0(a[0], b[0]), ccd::stitch0(a[1], b[1]), ccd::stitch0(a[2], b[2]));
        }
        double tt = 0.0;
        const int r = is_ee ? ccd::check_ee_window(s, e, eta, tt) : ccd::check_vf_window(s, e, eta, tt);
        hit[i] = r >= 0;
        t[i] = r >= 0 ? tt : 0.0;
        sub[i] = r;
    }
}

// The two-pass edge-edge schedule of the device (k_narrow2_queue_ee_shared with a deferred list, then
// k_ee_deferred), stencil by stencil on the host: first pass with ccd::ScreenSextic, the vertex-edge and
// vertex-vertex tests in the window's order for every stencil the edge-edge screen did not clear; a
// deferred stencil keeps that sub-test number, then gets the full edge-edge test and, if that misses,
// the recorded sub-test run again for its time.  Writes hit / t / sub as core_narrowphase does and
// deferred[i] = 1 for the stencils the first pass set aside; returns how many.
extern "C" int64_t core_ee_two_pass(const double *qs, const double *qe, const int32_t *keys, int64_t n, double eta,
                                    int32_t *hit, double *t, int32_t *sub, int32_t *deferred)
{
    int64_t nd = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nd)
    for (int64_t i = 0; i < n; i++) {
        V3 s[4], e[4];
        for (int j = 0; j < 4; j++) {
            const double *a = qs + 3 * (int64_t)keys[4 * i + j];
            const double *b = qe + 3 * (int64_t)keys[4 * i + j];
            e[j] = ld(b);
            s[j] = ccd::mk(ccd::stitch0(a[0], b[0]), ccd::stitch0(a[1], b[1]), ccd::stitch0(a[2], b[2]));
        }
        double tt = 0.0;
        int r = -1;
        ccd::ScreenSextic ec;   // the screen never reports a hit
        ccd::edge_edge(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, tt, ec);
        if (r < 0) {
            const int ve[4][3] = {{0, 2, 3}, {1, 2, 3}, {2, 0, 1}, {3, 0, 1}};
            for (int k = 0; k < 4 && r < 0; k++)
                if (ccd::vertex_edge(s[ve[k][0]], s[ve[k][1]], s[ve[k][2]], e[ve[k][0]], e[ve[k][1]], e[ve[k][2]], eta, tt))
                    r = 1 + k;
            const int vv[4][2] = {{0, 2}, {0, 3}, {1, 2}, {1, 3}};
            for (int k = 0; k < 4 && r < 0; k++)
                if (ccd::vertex_vertex(s[vv[k][0]], s[vv[k][1]], e[vv[k][0]], e[vv[k][1]], eta, tt)) r = 5 + k;
        }
        deferred[i] = ec.deferred ? 1 : 0;
        if (ec.deferred) {
            nd++;
            const int first = r;
            tt = 0.0;
            r = -1;
            if (ccd::edge_edge(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, tt)) {
                r = 0;
            } else if (first >= 1 && first <= 4) {
                const int k = first - 1, a = k < 2 ? 2 : 0;
                if (ccd::vertex_edge(s[k], s[a], s[a + 1], e[k], e[a], e[a + 1], eta, tt)) r = first;
            } else if (first >= 5) {
                const int x = first < 7 ? 0 : 1, y = (first == 5 || first == 7) ? 2 : 3;
                if (ccd::vertex_vertex(s[x], s[y], e[x], e[y], eta, tt)) r = first;
            }
        }
        hit[i] = r >= 0;
        t[i] = r >= 0 ? tt : 0.0;
        sub[i] = r;
    }
    return nd;
}

// findIntervals on a copy of op (degree 2, 3, 4 or 6); returns the count, pieces in out[2k], out[2k+1].
extern "C" int core_find_intervals(const double *op, int n, int pos, double *out)
{
    double c[7];
    for (int i = 0; i <= n; i++) c[i] = op[i];
    ccd::Ivls<8> L;
    if (n == 2) ccd::find_intervals<2>(c, L, pos != 0);
    else if (n == 3) ccd::find_intervals<3>(c, L, pos != 0);
    else if (n == 4) ccd::find_intervals<4>(c, L, pos != 0);
    else if (n == 6) ccd::find_intervals<6>(c, L, pos != 0);
    else return -1;
    int k = 0;
    for (unsigned m = L.keep; m; m &= m - 1, k++) L.get(ccd::low_bit(m), out[2 * k], out[2 * k + 1]);
    return k;
}

// The core's real roots of c[0..n] (n = 3..6) in [0,1].
extern "C" int core_roots01(const double *c, int n, double *r)
{
    double buf[6] = {0, 0, 0, 0, 0, 0};
    int k = -1;
    switch (n) {
    case 3: k = ccd::isolate_into<3>(c, buf); break;
    case 4: k = ccd::isolate_into<4>(c, buf); break;
    case 5: k = ccd::isolate_into<5>(c, buf); break;
    case 6: k = ccd::isolate_into<6>(c, buf); break;
    }
    for (int i = 0; i < n && i < 6; i++) r[i] = buf[i];
    return k;
}

// Proximity statics (kind as CCD_DIST_*): returns the flag for kinds 0/1, else writes vec[3], bary[4].
extern "C" int core_distance(int kind, const double *q, double eta, double *vec, double *bary)
{
    V3 a = ld(q), b = ld(q + 3), c = ld(q + 6), d = ld(q + 9);
    if (kind == 0) return ccd::vertex_plane_distance_less_than(a, b, c, d, eta);
    if (kind == 1) return ccd::line_line_distance_less_than(a, b, c, d, eta);
    V3 r = kind == 2 ? ccd::vertex_face_distance(a, b, c, d, bary[0], bary[1], bary[2])
                     : ccd::edge_edge_distance(a, b, c, d, bary[0], bary[1], bary[2], bary[3]);
    vec[0] = r.x;
    vec[1] = r.y;
    vec[2] = r.z;
    return 1;
}

// The separating-plane narrowphase of the core over a CSR history; eps = the minimum entry gap / 4
// (History::computeMinimumGap).  hit[i] = 1, 0, or 2 (undecided: the reference's recursion would not end).
extern "C" void core_separating_plane(const int64_t *off, const double *ht, const double *hp, int64_t nverts,
                                      const int32_t *keys, const double *eta, int64_t n, int is_ee, int32_t *hit)
{
    double gap = 1.0;
    for (int64_t v = 0; v < nverts; v++)
        for (int64_t j = off[v] + 1; j < off[v + 1]; j++) gap = ccd::smin(ht[j] - ht[j - 1], gap);
    const double eps = gap / 4.0;
    ccd::HistCSR h{off, ht, hp};
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        int v[4] = {keys[4 * i], keys[4 * i + 1], keys[4 * i + 2], keys[4 * i + 3]};
        if (g_sp_prune)
            hit[i] = is_ee ? ccd::separating_plane_check<false>(h, v, eta[i], eps)
                           : ccd::separating_plane_check<true>(h, v, eta[i], eps);
        else
            hit[i] = is_ee ? ccd::separating_plane_check<false>(h, v, eta[i], eps, ccd::FullCascade())
                           : ccd::separating_plane_check<true>(h, v, eta[i], eps, ccd::FullCascade());
    }
}
