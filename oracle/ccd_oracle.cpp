// TEST INFRASTRUCTURE ONLY -- the CPU oracle.
//
// A CPU restatement of evouga/collisiondetection's swept-primitive CCD path, written from a reading of
// the reference's behaviour.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker (or the timed CPU baseline); the product never does.
//
// Every function names the reference file:line it follows.  Floating-point operations are written in
// the reference's operation order and must be compiled with -ffp-contract=off (see oracle/Makefile):
//   * Eigen's Vector3d::dot/squaredNorm are summed ((x0*y0 + x1*y1) + x2*y2), the order Eigen >= 3.3
//     produces for a 3-vector with SSE2 packets (the reference Makefile builds -O2 with no -march);
//     orc_set_dot_order(1) switches to x0*y0 + (x1*y1 + x2*y2) (Eigen 3.2's unvectorised unroller).
//   * Eigen's cross() is (a1*b2 - a2*b1, a2*b0 - a0*b2, a0*b1 - a1*b0).
//   * std::max(a,b) is (a < b) ? b : a and std::min(a,b) is (b < a) ? b : a, argument order kept.
//
// The reference's polynomial root finder (src/rpoly.h, Jenkins-Traub) is not restated.  By default the
// oracle finds all complex roots with its own Aberth-Ehrlich iteration in long double and, like rpoly,
// hands back the real parts of all of them.  orc_set_rootfinder() substitutes any other function with
// rpoly's signature -- the tests use it to plug in the reference's own rpoly built in place under
// oracle/_ref/ from /root/reference/src/rpoly.h.
//
// Besides the reference's results, the oracle reports for every primitive test a decision margin: the
// smallest distance between two quantities whose comparison decided the result (a root and 0 or 1, two
// roots, two interval endpoints from different lists).  A result with a margin below 1e-12 is
// "borderline": a different but equally accurate root finder may flip it.

#include <algorithm>
#include <cmath>
#include <complex>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <queue>
#include <set>
#include <vector>

namespace {

// ----------------------------------------------------------------------------------------------
// Globals: dot order and root finder.
// ----------------------------------------------------------------------------------------------
int g_dot_order = 0;
typedef int (*rootfinder_fn)(double *op, int degree, double *zeror, double *zeroi);
rootfinder_fn g_rootfinder = nullptr;   // nullptr = built-in Aberth

// std::max / std::min with the reference's argument order semantics.
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline double smin(double a, double b) { return (b < a) ? b : a; }

struct V3 {
    double x, y, z;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b)
{
    if (g_dot_order == 0) return (a.x * b.x + a.y * b.y) + a.z * b.z;
    return a.x * b.x + (a.y * b.y + a.z * b.z);
}
inline V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// ----------------------------------------------------------------------------------------------
// Work counters (for the roofline's algorithmic flop count) and the decision margin.
// ----------------------------------------------------------------------------------------------
enum Stage {
    ST_VF_CALL = 0,    // one vertexFaceCTCD call (CTCD.cpp:413-427)
    ST_VF_PLANE,       // one edge-side cubic of vertexFaceCTCD (CTCD.cpp:433-464)
    ST_VF_DIST,        // the coplanar-distance sextic of vertexFaceCTCD (CTCD.cpp:466-475)
    ST_EE_DIST,        // the distance sextic of edgeEdgeCTCD (CTCD.cpp:275-288)
    ST_EE_PAR,         // one parallel-edge check of a raw coplanar interval (CTCD.cpp:294-308)
    ST_EE_BARY,        // one barycentric quartic of edgeEdgeCTCD (CTCD.cpp:313-350)
    ST_VE_CALL,        // one vertexEdgeCTCD call (CTCD.cpp:511-534)
    ST_VE_QUAD,        // one side quadratic of vertexEdgeCTCD (CTCD.cpp:529-555)
    ST_VE_QUART,       // the colinear-distance quartic of vertexEdgeCTCD (CTCD.cpp:557-573)
    ST_VV,             // vertexVertexCTCD's quadratic (CTCD.cpp:604-639)
    ST_QUAD_ROOT,      // one getQuadRoots call that found roots (CTCD.cpp:38-56)
    ST_ISOLATE,        // one polynomial of degree > 2 handed to the root finder (CTCD.cpp:135-142)
    ST_COUNT
};

struct Ctx {
    double margin = std::numeric_limits<double>::infinity();
    bool complex_split = false;   // a complex root's real part split [0,1] in the last find_intervals
    int64_t counts[ST_COUNT] = {0};
    void gap(double a, double b)
    {
        double d = std::fabs(a - b);
        if (d < margin) margin = d;
    }
    // A comparison between two interval endpoints.  Two endpoints that are both exactly 0 or both
    // exactly 1 come from clamping (or from an exact zero root) and compare the same under any root
    // finder, so they do not count.
    // A comparison counts when one end is a root-finder root and the other any real root; a root
    // against a clamped end (0 or 1) is counted where the root is found.
    void endpoint_gap(double a, int ak, double b, int bk)
    {
        if ((ak == 2 && (bk == 1 || bk == 2)) || (bk == 2 && ak == 1)) gap(a, b);
    }
};

// ----------------------------------------------------------------------------------------------
// Time intervals: include/CTCD.h:7-26, src/CTCD.cpp:10-36
// ----------------------------------------------------------------------------------------------
struct TI {
    double l, u;
    // Provenance of each end, for the decision margin only: 0 = clamp or constant, 1 = a root from
    // a closed form (degree <= 2, identical under any root finder), 2 = a real root found by the
    // root finder (degree > 2), 3 = the real part of a complex root.
    int lk = 0, uk = 0;
};
inline TI make_ti(double tl, double tu)   // CTCD.h:9-14
{
    TI r;
    r.l = tl;
    r.u = tu;
    if (r.l > r.u) std::swap(r.l, r.u);
    r.l = smax(r.l, 0.0);
    r.u = smin(r.u, 1.0);
    return r;
}
inline bool ti_overlap(const TI &a, const TI &b) { return !(a.l > b.u || b.l > a.u); }   // CTCD.cpp:10-13
bool ti_overlap_all(const TI *iv, int n, Ctx *ctx)   // CTCD.cpp:15-25 (pairwise)
{
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            if (ctx) {
                ctx->endpoint_gap(iv[i].l, iv[i].lk, iv[j].u, iv[j].uk);
                ctx->endpoint_gap(iv[j].l, iv[j].lk, iv[i].u, iv[i].uk);
            }
            if (!ti_overlap(iv[i], iv[j])) return false;
        }
    return true;
}
TI ti_intersect(const TI *iv, int n)   // CTCD.cpp:27-36
{
    TI isect = make_ti(0.0, 1.0);
    for (int i = 0; i < n; i++) {
        isect.l = smax(iv[i].l, isect.l);
        isect.u = smin(iv[i].u, isect.u);
    }
    return isect;
}

// ----------------------------------------------------------------------------------------------
// Root finding.
// ----------------------------------------------------------------------------------------------

// Built-in: Aberth-Ehrlich simultaneous iteration (O. Aberth 1973, L. W. Ehrlich 1967) on the
// polynomial with descending coefficients op[0..degree], in long double complex arithmetic.  Like
// rpoly it returns the real parts of all roots and strips exact zeros at the origin first, so a root
// that is exactly 0 is reported as exactly 0.0.
int aberth_roots(const double *op, int degree, double *zeror, double *zeroi)
{
    typedef long double LD;
    typedef std::complex<LD> C;
    if (op[0] == 0.0) return -1;
    int n = degree;
    int nz = 0;
    while (n > 0 && op[n] == 0.0) {
        zeror[nz] = 0.0;
        zeroi[nz] = 0.0;
        nz++;
        n--;
    }
    if (n < 1) return nz;
    LD a[8];
    for (int i = 0; i <= n; i++) a[i] = (LD)op[i] / (LD)op[0];
    if (n == 1) {
        zeror[nz] = (double)(-a[1]);
        zeroi[nz] = 0.0;
        return nz + 1;
    }
    // Initial guesses on a circle whose radius is the Fujiwara bound, rotated off the real axis.
    LD rad = 0;
    for (int i = 1; i <= n; i++) {
        LD c = std::fabs(a[i]);
        if (i == n) c /= 2;
        LD r = std::pow(c, (LD)1.0 / (LD)i);
        if (r > rad) rad = r;
    }
    rad *= 2;
    if (!(rad > 0)) rad = 1;
    C z[8];
    const LD two_pi = 6.283185307179586476925286766559L;
    for (int k = 0; k < n; k++) z[k] = std::polar(rad, two_pi * k / n + 0.4L);
    bool done[8] = {false};
    LD lastw[8];
    for (int k = 0; k < n; k++) lastw[k] = std::numeric_limits<LD>::infinity();
    for (int it = 0; it < 500; it++) {
        bool all = true;
        for (int k = 0; k < n; k++) {
            if (done[k]) continue;
            C p = a[0], dp = 0;
            for (int i = 1; i <= n; i++) {
                dp = dp * z[k] + p;
                p = p * z[k] + a[i];
            }
            if (p == C(0)) {
                done[k] = true;
                continue;
            }
            C ratio = p / dp;
            C s = 0;
            for (int j = 0; j < n; j++)
                if (j != k) s += C(1) / (z[k] - z[j]);
            C w = ratio / (C(1) - ratio * s);
            if (!std::isfinite(std::real(w)) || !std::isfinite(std::imag(w))) {
                w = ratio;
                if (!std::isfinite(std::real(w)) || !std::isfinite(std::imag(w))) w = C(1e-9L, 1e-9L);
            }
            z[k] -= w;
            LD aw = std::abs(w), sc = std::max(std::abs(z[k]), (LD)1);
            // Converged at working precision, or stalled at the rounding-noise floor (a clustered
            // or multiple root, whose correction stops shrinking).
            if (aw <= 1e-18L * sc || (it > 30 && aw < 1e-6L * sc && aw >= 0.5L * lastw[k]))
                done[k] = true;
            else
                all = false;
            lastw[k] = aw;
        }
        if (all) break;
    }
    // Newton polish of each root on the polynomial itself (damps the linear convergence that
    // Aberth shows at multiple roots only a little, but sharpens simple roots to long double).
    for (int k = 0; k < n; k++) {
        for (int it = 0; it < 3; it++) {
            C p = a[0], dp = 0;
            for (int i = 1; i <= n; i++) {
                dp = dp * z[k] + p;
                p = p * z[k] + a[i];
            }
            if (dp == C(0)) break;
            C w = p / dp;
            if (std::abs(w) > 1e-12L * std::max(std::abs(z[k]), (LD)1)) break;
            z[k] -= w;
        }
        zeror[nz + k] = (double)std::real(z[k]);
        zeroi[nz + k] = (double)std::imag(z[k]);
    }
    return nz + n;
}

int find_roots(double *op, int degree, double *zr, double *zi)
{
    if (g_rootfinder) {
        double tmp[8];
        for (int i = 0; i <= degree; i++) tmp[i] = op[i];
        return g_rootfinder(tmp, degree, zr, zi);
    }
    return aberth_roots(op, degree, zr, zi);
}

// ----------------------------------------------------------------------------------------------
// Sign intervals of a polynomial: src/CTCD.cpp:38-177
// ----------------------------------------------------------------------------------------------

// CTCD.cpp:38-56
int get_quad_roots(double a, double b, double c, double &t0, double &t1)
{
    int roots = 0;
    int sign = 1;
    if (b < 0) sign = -1;
    double D = b * b - 4 * a * c;
    if (D >= 0) {
        roots = 2;
        double q = -0.5 * (b + sign * std::sqrt(D));
        t0 = q / a;
        t1 = c / q;
        if (t0 > t1) std::swap(t1, t0);
    }
    return roots;
}

// CTCD.cpp:59-79
void check_interval(double t1, double t2, const double *op, int degree, std::vector<TI> &out, bool pos,
                    int r1 = 0, int r2 = 0)
{
    double c1 = smax(0.0, t1);
    double c2 = smax(0.0, t2);
    c1 = smin(1.0, c1);
    c2 = smin(1.0, c2);
    if (c1 != t1) r1 = 0;   // a clamped end is no longer a root
    if (c2 != t2) r2 = 0;
    t1 = c1;
    t2 = c2;
    double tmid = (t2 + t1) / 2;
    double f = op[0];
    for (int i = 1; i <= degree; i++) {
        f *= tmid;
        f += op[i];
    }
    if ((pos && f >= 0) || (!pos && f <= 0)) {
        TI ti = make_ti(t1, t2);
        ti.lk = (t1 > t2) ? r2 : r1;
        ti.uk = (t1 > t2) ? r1 : r2;
        out.push_back(ti);
    }
}

// CTCD.cpp:81-94
bool could_have_roots(const double *op, int degree, bool pos)
{
    double result = 0;
    if ((pos && op[0] > 0) || (!pos && op[0] < 0)) result = op[0];
    for (int i = 1; i < degree; i++) {
        result *= 1.0;
        if ((pos && op[i] > 0) || (!pos && op[i] < 0)) result += op[i];
    }
    result *= 1.0;
    result += op[degree];
    return !((pos && result < 0) || (!pos && result > 0));
}

// CTCD.cpp:98-177.  op is modified in place (normalised and degree-reduced) as in the reference.
void find_intervals(double *op, int n, std::vector<TI> &out, bool pos, Ctx *ctx)
{
    int roots = 0;
    int reduced = n;
    if (n > 6) return;
    double time[8];
    double zeroi[8];
    double maxval = 0;
    for (int i = 0; i <= n; i++) maxval = smax(maxval, std::fabs(op[i]));
    if (maxval != 0)
        for (int i = 0; i <= n; i++) op[i] /= maxval;
    for (int i = 0; i < n; i++) {
        if (op[i] == 0)
            reduced--;
        else
            break;
    }
    if (reduced < n)
        for (int i = 0; i <= reduced; i++) op[i] = op[i + n - reduced];

    if (reduced > 2) {
        if (!could_have_roots(op, reduced, pos)) return;
        if (ctx) ctx->counts[ST_ISOLATE]++;
        roots = find_roots(op, reduced, time, zeroi);
    } else if (reduced == 2) {
        roots = get_quad_roots(op[0], op[1], op[2], time[0], time[1]);
        if (ctx && roots > 0) ctx->counts[ST_QUAD_ROOT]++;
    } else if (reduced == 1) {
        time[0] = -op[1] / op[0];
        roots = 1;
    } else {
        if ((!pos && op[0] <= 0) || (pos && op[0] >= 0)) out.push_back(make_ti(0, 1.0));
        return;
    }

    if (roots > 0) {
        double zr_unsorted[8];
        for (int i = 0; i < roots; i++) zr_unsorted[i] = time[i];
        if (reduced <= 2)
            for (int i = 0; i < roots; i++) zeroi[i] = 0.0;
        std::sort(time, time + roots);
        if (ctx && reduced > 2) {
            // Margin: real roots (imaginary part below 1e-6) that split [0,1], against the interval
            // ends and against each other.  The real part of a complex pair also splits the
            // interval in the reference, but the polynomial keeps its sign across it, so the union
            // of the intervals -- all that vertex-face and vertex-edge tests use -- does not depend
            // on it.  Edge-edge tests classify each piece as parallel or not, which can depend on
            // it; that case is flagged in edge_edge_ctcd through ctx->complex_split.
            double rr[8];
            int nr = 0;
            for (int i = 0; i < roots; i++) {
                bool real = std::fabs(zeroi[i]) < 1e-6;
                if (real)
                    rr[nr++] = zr_unsorted[i];
                else if (zr_unsorted[i] >= 0.0 && zr_unsorted[i] <= 1.0)
                    ctx->complex_split = true;
            }
            std::sort(rr, rr + nr);
            for (int i = 0; i < nr; i++) {
                if (rr[i] >= -1e-6 && rr[i] <= 1.0 + 1e-6) {
                    if (!(rr[i] == 0.0 && op[reduced] == 0.0)) ctx->gap(rr[i], 0.0);
                    ctx->gap(rr[i], 1.0);
                    if (i + 1 < nr) ctx->gap(rr[i], rr[i + 1]);
                }
            }
        }
        int isreal[8];
        for (int i = 0; i < roots; i++) {
            isreal[i] = reduced > 2 ? 2 : 1;
            for (int j = 0; j < roots; j++)
                if (zr_unsorted[j] == time[i] && !(std::fabs(zeroi[j]) < 1e-6)) isreal[i] = 3;
        }
        if (time[0] >= 0) check_interval(0, time[0], op, reduced, out, pos, 0, isreal[0]);
        for (int i = 0; i < roots - 1; i++) {
            if (!((time[i] < 0 && time[i + 1] < 0) || (time[i] > 1.0 && time[i + 1] > 1.0)))
                check_interval(time[i], time[i + 1], op, reduced, out, pos, isreal[i], isreal[i + 1]);
        }
        if (time[roots - 1] <= 1.0)
            check_interval(time[roots - 1], 1.0, op, reduced, out, pos, isreal[roots - 1], 0);
    } else {
        check_interval(0.0, 1.0, op, reduced, out, pos);
    }
}

// CTCD.cpp:179-213
void barycentric_poly(V3 x10, V3 x20, V3 x30, V3 v10, V3 v20, V3 v30, std::vector<TI> &out, Ctx *ctx)
{
    if (ctx) ctx->counts[ST_EE_BARY]++;
    double A = dot(x10, x10);
    double B = 2 * dot(x10, v10);
    double C = dot(v10, v10);
    double D = dot(x20, x10);
    double E = dot(x20, v10) + dot(v20, x10);
    double F = dot(v20, v10);
    double G = dot(x30, x20);
    double H = dot(x30, v20) + dot(v30, x20);
    double I = dot(v30, v20);
    double J = dot(x30, x10);
    double K = dot(x30, v10) + dot(v30, x10);
    double L = dot(v30, v10);
    double op[5];
    op[0] = F * L - C * I;
    op[1] = F * K + E * L - C * H - B * I;
    op[2] = F * J + D * L + E * K - C * G - A * I - B * H;
    op[3] = D * K + E * J - A * H - B * G;
    op[4] = D * J - A * G;
    find_intervals(op, 4, out, true, ctx);
}

// CTCD.cpp:215-229
void plane_poly(V3 x10, V3 x20, V3 x30, V3 v10, V3 v20, V3 v30, std::vector<TI> &out, Ctx *ctx)
{
    if (ctx) ctx->counts[ST_VF_PLANE]++;
    double op[4];
    op[0] = dot(v10, cross(v20, v30));
    op[1] = dot(x10, cross(v20, v30)) + dot(v10, cross(x20, v30)) + dot(v10, cross(v20, x30));
    op[2] = dot(x10, cross(x20, v30)) + dot(x10, cross(v20, x30)) + dot(v10, cross(x20, x30));
    op[3] = dot(x10, cross(x20, x30));
    find_intervals(op, 3, out, true, ctx);
}

// CTCD.cpp:231-257
void distance_poly(V3 x10, V3 x20, V3 x30, V3 v10, V3 v20, V3 v30, double minD, std::vector<TI> &out,
                   Ctx *ctx)
{
    double A = dot(v10, cross(v20, v30));
    double B = dot(x10, cross(v20, v30)) + dot(v10, cross(x20, v30)) + dot(v10, cross(v20, x30));
    double C = dot(x10, cross(x20, v30)) + dot(x10, cross(v20, x30)) + dot(v10, cross(x20, x30));
    double D = dot(x10, cross(x20, x30));
    V3 E = cross(x20, x30);
    V3 F = cross(x20, v30) + cross(v20, x30);
    V3 G = cross(v20, v30);
    double op[7];
    op[0] = A * A;
    op[1] = 2 * A * B;
    op[2] = B * B + 2 * A * C - dot(G, G) * minD;
    op[3] = 2 * A * D + 2 * B * C - 2 * dot(G, F) * minD;
    op[4] = 2 * B * D + C * C - (2 * dot(G, E) + dot(F, F)) * minD;
    op[5] = 2 * C * D - 2 * dot(F, E) * minD;
    op[6] = D * D - dot(E, E) * minD;
    find_intervals(op, 6, out, false, ctx);
}

// ----------------------------------------------------------------------------------------------
// Primitive tests: src/CTCD.cpp:259-692
// ----------------------------------------------------------------------------------------------

// CTCD.cpp:259-411.  Edges (q0,p0) and (q1,p1).
bool edge_edge_ctcd(V3 q0s, V3 p0s, V3 q1s, V3 p1s, V3 q0e, V3 p0e, V3 q1e, V3 p1e, double eta, double &t,
                    Ctx *ctx)
{
    double minD = eta * eta;
    std::vector<TI> rawcoplane, a0, a1, b0, b1;
    V3 x10 = p0s - p1s;
    V3 x20 = p0s - q0s;
    V3 x30 = p1s - q1s;
    V3 vp0 = p0e - p0s;
    V3 vp1 = p1e - p1s;
    V3 vq0 = q0e - q0s;
    V3 vq1 = q1e - q1s;
    V3 v10 = vp0 - vp1;
    V3 v20 = vp0 - vq0;
    V3 v30 = vp1 - vq1;
    if (ctx) {
        ctx->counts[ST_EE_DIST]++;
        ctx->complex_split = false;
    }
    distance_poly(x10, x20, x30, v10, v20, v30, minD, rawcoplane, ctx);
    bool raw_complex_split = ctx && ctx->complex_split;

    std::vector<TI> coplane, parallel;
    for (size_t i = 0; i < rawcoplane.size(); i++) {
        if (ctx) ctx->counts[ST_EE_PAR]++;
        double midt = (rawcoplane[i].u + rawcoplane[i].l) / 2;
        x10 = (q0s - p0s) + midt * (vq0 - vp0);
        x20 = (q1s - p1s) + midt * (vq1 - vp1);
        double cn = norm(cross(x10, x20));
        if (ctx) ctx->gap(cn / 1e-8, 1.0);   // relative distance to the absolute 1e-8 threshold
        if (cn < 1e-8)
            parallel.push_back(rawcoplane[i]);
        else
            coplane.push_back(rawcoplane[i]);
    }
    if (coplane.empty()) return false;
    // Pieces split at a complex root's real part are classified parallel or not one by one, so with
    // a parallel piece the result can depend on those split points: flag it.
    if (ctx && raw_complex_split && !parallel.empty()) ctx->margin = 0.0;

    x10 = p1s - q1s;
    v10 = vp1 - vq1;
    x20 = p0s - q0s;
    v20 = vp0 - vq0;
    x30 = q0s - q1s;
    v30 = vq0 - vq1;
    barycentric_poly(x10, x20, x30, v10, v20, v30, a0, ctx);
    if (a0.empty()) return false;

    x20 = q0s - p0s;
    v20 = vq0 - vp0;
    x30 = p0s - q1s;
    v30 = vp0 - vq1;
    barycentric_poly(x10, x20, x30, v10, v20, v30, a1, ctx);
    if (a1.empty()) return false;

    x10 = p0s - q0s;
    v10 = vp0 - vq0;
    x20 = p1s - q1s;
    v20 = vp1 - vq1;
    x30 = q1s - q0s;
    v30 = vq1 - vq0;
    barycentric_poly(x10, x20, x30, v10, v20, v30, b0, ctx);
    if (b0.empty()) return false;

    x20 = q1s - p1s;
    v20 = vq1 - vp1;
    x30 = p1s - q0s;
    v30 = vp1 - vq0;
    barycentric_poly(x10, x20, x30, v10, v20, v30, b1, ctx);
    if (b1.empty()) return false;

    bool col = false;
    double mint = 1.0;
    TI iv[5];
    for (size_t i = 0; i < coplane.size(); i++)
        for (size_t j = 0; j < a0.size(); j++)
            for (size_t k = 0; k < a1.size(); k++)
                for (size_t l = 0; l < b0.size(); l++)
                    for (size_t m = 0; m < b1.size(); m++) {
                        iv[0] = coplane[i];
                        iv[1] = a0[j];
                        iv[2] = a1[k];
                        iv[3] = b0[l];
                        iv[4] = b1[m];
                        if (ti_overlap_all(iv, 5, ctx)) {
                            TI isect = ti_intersect(iv, 5);
                            bool skip = false;
                            for (size_t p = 0; p < parallel.size(); p++) {
                                TI pc[2] = {isect, parallel[p]};
                                if (ti_overlap_all(pc, 2, ctx)) {
                                    skip = true;
                                    break;
                                }
                            }
                            if (!skip) {
                                mint = smin(mint, isect.l);
                                col = true;
                            }
                        }
                    }
    if (col) {
        t = mint;
        return true;
    }
    return false;
}

// CTCD.cpp:413-508.  Vertex q0 against face (q1,q2,q3).
bool vertex_face_ctcd(V3 q0s, V3 q1s, V3 q2s, V3 q3s, V3 q0e, V3 q1e, V3 q2e, V3 q3e, double eta, double &t,
                      Ctx *ctx)
{
    if (ctx) ctx->counts[ST_VF_CALL]++;
    double minD = eta * eta;
    V3 v0 = q0e - q0s;
    V3 v1 = q1e - q1s;
    V3 v2 = q2e - q2s;
    V3 v3_ = q3e - q3s;
    std::vector<TI> coplane, e1, e2, e3;

    V3 x10 = q0s - q1s;
    V3 v10 = v0 - v1;
    V3 x20 = cross(q3s - q1s, q2s - q1s);
    V3 v20 = cross(v3_ - v1, v2 - v1);
    V3 x30 = q3s - q1s;
    V3 v30 = v3_ - v1;
    plane_poly(x10, x20, x30, v10, v20, v30, e1, ctx);
    if (e1.empty()) return false;

    x10 = q0s - q2s;
    v10 = v0 - v2;
    x20 = cross(q1s - q2s, q3s - q2s);
    v20 = cross(v1 - v2, v3_ - v2);
    x30 = q1s - q2s;
    v30 = v1 - v2;
    plane_poly(x10, x20, x30, v10, v20, v30, e2, ctx);
    if (e2.empty()) return false;

    x10 = q0s - q3s;
    v10 = v0 - v3_;
    x20 = cross(q2s - q3s, q1s - q3s);
    v20 = cross(v2 - v3_, v1 - v3_);
    x30 = q2s - q3s;
    v30 = v2 - v3_;
    plane_poly(x10, x20, x30, v10, v20, v30, e3, ctx);
    if (e3.empty()) return false;

    x10 = q0s - q1s;
    x20 = q2s - q1s;
    x30 = q3s - q1s;
    v10 = v0 - v1;
    v20 = v2 - v1;
    v30 = v3_ - v1;
    if (ctx) ctx->counts[ST_VF_DIST]++;
    distance_poly(x10, x20, x30, v10, v20, v30, minD, coplane, ctx);
    if (coplane.empty()) return false;

    bool col = false;
    double mint = 1.0;
    TI iv[4];
    for (size_t i = 0; i < coplane.size(); i++)
        for (size_t j = 0; j < e1.size(); j++)
            for (size_t k = 0; k < e2.size(); k++)
                for (size_t l = 0; l < e3.size(); l++) {
                    iv[0] = coplane[i];
                    iv[1] = e1[j];
                    iv[2] = e2[k];
                    iv[3] = e3[l];
                    if (ti_overlap_all(iv, 4, ctx)) {
                        mint = smin(ti_intersect(iv, 4).l, mint);
                        col = true;
                    }
                }
    if (col) {
        t = mint;
        return true;
    }
    return false;
}

// CTCD.cpp:511-602.  Vertex q0 against edge (q1,q2).
bool vertex_edge_ctcd(V3 q0s, V3 q1s, V3 q2s, V3 q0e, V3 q1e, V3 q2e, double eta, double &t, Ctx *ctx)
{
    if (ctx) ctx->counts[ST_VE_CALL]++;
    double op[5];
    double minD = eta * eta;
    V3 v0 = q0e - q0s;
    V3 v1 = q1e - q1s;
    V3 v2 = q2e - q2s;
    std::vector<TI> colin, e1, e2;
    V3 ab = q2s - q1s;
    V3 ac = q0s - q1s;
    V3 cb = q2s - q0s;
    V3 vab = v2 - v1;
    V3 vac = v0 - v1;
    V3 vcb = v2 - v0;

    if (ctx) ctx->counts[ST_VE_QUAD]++;
    double c = dot(ab, ac);
    double b = dot(ac, vab) + dot(ab, vac);
    double a = dot(vab, vac);
    op[0] = a;
    op[1] = b;
    op[2] = c;
    find_intervals(op, 2, e1, true, ctx);
    if (e1.empty()) return false;

    if (ctx) ctx->counts[ST_VE_QUAD]++;
    c = dot(ab, cb);
    b = dot(cb, vab) + dot(ab, vcb);
    a = dot(vab, vcb);
    op[0] = a;
    op[1] = b;
    op[2] = c;
    find_intervals(op, 2, e2, true, ctx);
    if (e2.empty()) return false;

    if (ctx) ctx->counts[ST_VE_QUART]++;
    double A = dot(ab, ab);
    double B = 2 * dot(ab, vab);
    double C = dot(vab, vab);
    double D = dot(ac, ac);
    double E = 2 * dot(ac, vac);
    double F = dot(vac, vac);
    double G = dot(ac, ab);
    double H = dot(vab, ac) + dot(vac, ab);
    double I = dot(vab, vac);
    op[4] = A * D - G * G - minD * A;
    op[3] = B * D + A * E - 2 * G * H - minD * B;
    op[2] = B * E + A * F + C * D - H * H - 2 * G * I - minD * C;
    op[1] = B * F + C * E - 2 * H * I;
    op[0] = C * F - I * I;
    find_intervals(op, 4, colin, false, ctx);
    if (colin.empty()) return false;

    double mint = 1.0;
    bool col = false;
    TI iv[3];
    for (size_t i = 0; i < colin.size(); i++)
        for (size_t j = 0; j < e1.size(); j++)
            for (size_t k = 0; k < e2.size(); k++) {
                iv[0] = colin[i];
                iv[1] = e1[j];
                iv[2] = e2[k];
                if (ti_overlap_all(iv, 3, ctx)) {
                    mint = smin(ti_intersect(iv, 3).l, mint);
                    col = true;
                }
            }
    if (col) {
        t = mint;
        return true;
    }
    return false;
}

// CTCD.cpp:604-692
bool vertex_vertex_ctcd(V3 q1s, V3 q2s, V3 q1e, V3 q2e, double eta, double &t, Ctx *ctx)
{
    if (ctx) ctx->counts[ST_VV]++;
    int roots = 0;
    double min_d = eta * eta;
    double t1 = 0, t2 = 0;
    V3 v1 = q1e - q1s;
    V3 v2 = q2e - q2s;
    double a = dot(v1, v1) + dot(v2, v2) - 2 * dot(v1, v2);
    double b = 2 * (dot(v1, q1s) - dot(v2, q1s) - dot(v1, q2s) + dot(v2, q2s));
    double c = dot(q1s, q1s) + dot(q2s, q2s) - 2 * dot(q1s, q2s) - min_d;
    if (a != 0) {
        roots = get_quad_roots(a, b, c, t1, t2);
        if (ctx && roots > 0) ctx->counts[ST_QUAD_ROOT]++;
    } else if (b != 0) {
        t1 = -c / b;
        roots = 1;
    } else {
        if (c <= 0) {
            t = 0;
            return true;
        }
        return false;
    }
    double op[3] = {a, b, c};
    std::vector<TI> iv;
    if (roots == 2) {
        check_interval(0, t1, op, 2, iv, false);
        if (!iv.empty()) {
            t = 0;
            return true;
        }
        check_interval(t1, t2, op, 2, iv, false);
        if (!iv.empty()) {
            t = t1;
            return true;
        }
        check_interval(t2, 1.0, op, 2, iv, false);
        if (!iv.empty()) {
            t = t2;
            return true;
        }
        return false;
    } else if (roots == 1) {
        check_interval(0, t1, op, 2, iv, false);
        if (!iv.empty()) {
            t = 0;
            return true;
        }
        check_interval(t1, 1.0, op, 2, iv, false);
        if (!iv.empty()) {
            t = t1;
            return true;
        }
        return false;
    }
    check_interval(0, 1.0, op, 2, iv, false);
    if (!iv.empty()) {
        t = 0;
        return true;
    }
    return false;
}

inline V3 load3(const double *p) { return V3{p[0], p[1], p[2]}; }

// ----------------------------------------------------------------------------------------------
// Histories: src/History.cpp:8-39 (construction) and :98-140 (stitchCommonHistory).
// A history is CSR: vertex v owns entries off[v] .. off[v+1]-1 with times ht[] and positions hp[3*].
// ----------------------------------------------------------------------------------------------
struct Hist {
    const int64_t *off;
    const double *ht;
    const double *hp;
};

// History.cpp:98-140 for nv vertices.  Produces stitched times and nv positions per entry.
void stitch(const Hist &h, const int *verts, int nv, std::vector<double> &times, std::vector<V3> &pos)
{
    times.clear();
    pos.clear();
    int64_t its[4];
    for (int i = 0; i < nv; i++) its[i] = h.off[verts[i]];
    double curtime = 0;
    while (curtime <= 1.0) {
        double newtime = std::numeric_limits<double>::infinity();
        times.push_back(curtime);
        for (int i = 0; i < nv; i++) {
            int64_t end = h.off[verts[i] + 1];
            int64_t next = its[i] + 1;
            while (next != end && h.ht[next] <= curtime) {
                ++next;
                ++its[i];
            }
            V3 oldpos = load3(h.hp + 3 * its[i]);
            if (next == end)
                pos.push_back(oldpos);
            else {
                V3 newpos = load3(h.hp + 3 * next);
                double dt = h.ht[next] - h.ht[its[i]];
                double a = curtime - h.ht[its[i]];
                double b = h.ht[next] - curtime;
                V3 num = b * oldpos + a * newpos;
                pos.push_back(V3{num.x / dt, num.y / dt, num.z / dt});
                newtime = smin(newtime, h.ht[next]);
            }
        }
        curtime = newtime;
    }
}

// Per-stencil result record shared with the Python side (tests/ read it through ctypes).
struct StencilResult {
    int32_t hit;       // 1 if the reference's narrowphase reports the stencil
    int32_t subtest;   // which primitive fired: 0 = VF/EE, 1.. = VE in reference order, then VV
    int32_t window;    // stitched window index of the hit
    int32_t pad;
    double t;          // the firing primitive's time of impact (window-local, as the reference's t)
    double margin;     // decision margin over every primitive executed for this stencil
};

// CTCDNarrowPhase.cpp:24-72
bool check_vfs(const Hist &h, const int *key, double eta, StencilResult &res, Ctx *ctx)
{
    int verts[4] = {key[0], key[1], key[2], key[3]};
    std::vector<double> times;
    std::vector<V3> P;
    stitch(h, verts, 4, times, P);
    size_t ne = times.size();
    for (size_t w = 0; w + 1 < ne; w++) {
        const V3 *s = &P[4 * w];
        const V3 *e = &P[4 * (w + 1)];
        double t;
        res.window = (int)w;
        if (vertex_face_ctcd(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, t, ctx)) {
            res.subtest = 0;
            res.t = t;
            return true;
        }
        for (int edge = 0; edge < 3; edge++) {
            int a = 1 + (edge % 3), b = 1 + ((edge + 1) % 3);
            if (vertex_edge_ctcd(s[0], s[a], s[b], e[0], e[a], e[b], eta, t, ctx)) {
                res.subtest = 1 + edge;
                res.t = t;
                return true;
            }
        }
        for (int v = 0; v < 3; v++) {
            if (vertex_vertex_ctcd(s[0], s[1 + v], e[0], e[1 + v], eta, t, ctx)) {
                res.subtest = 4 + v;
                res.t = t;
                return true;
            }
        }
    }
    return false;
}

// CTCDNarrowPhase.cpp:74-135
bool check_ees(const Hist &h, const int *key, double eta, StencilResult &res, Ctx *ctx)
{
    int verts[4] = {key[0], key[1], key[2], key[3]};
    std::vector<double> times;
    std::vector<V3> P;
    stitch(h, verts, 4, times, P);
    size_t ne = times.size();
    static const int ve_idx[4][3] = {{0, 2, 3}, {1, 2, 3}, {2, 0, 1}, {3, 0, 1}};
    static const int vv_idx[4][2] = {{0, 2}, {0, 3}, {1, 2}, {1, 3}};
    for (size_t w = 0; w + 1 < ne; w++) {
        const V3 *s = &P[4 * w];
        const V3 *e = &P[4 * (w + 1)];
        double t;
        res.window = (int)w;
        if (edge_edge_ctcd(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, t, ctx)) {
            res.subtest = 0;
            res.t = t;
            return true;
        }
        for (int k = 0; k < 4; k++) {
            const int *q = ve_idx[k];
            if (vertex_edge_ctcd(s[q[0]], s[q[1]], s[q[2]], e[q[0]], e[q[1]], e[q[2]], eta, t, ctx)) {
                res.subtest = 1 + k;
                res.t = t;
                return true;
            }
        }
        for (int k = 0; k < 4; k++) {
            const int *q = vv_idx[k];
            if (vertex_vertex_ctcd(s[q[0]], s[q[1]], e[q[0]], e[q[1]], eta, t, ctx)) {
                res.subtest = 5 + k;
                res.t = t;
                return true;
            }
        }
    }
    return false;
}

// ----------------------------------------------------------------------------------------------
// Stencil keys: src/Stencils.h:13-57 (VF) and :86-117 (EE)
// ----------------------------------------------------------------------------------------------
struct Key {
    int32_t k[4];
    bool operator<(const Key &o) const
    {
        for (int i = 0; i < 4; i++) {
            if (k[i] < o.k[i]) return true;
            if (k[i] > o.k[i]) return false;
        }
        return false;
    }
    bool operator==(const Key &o) const
    {
        return k[0] == o.k[0] && k[1] == o.k[1] && k[2] == o.k[2] && k[3] == o.k[3];
    }
};

Key vf_key(int p, int a, int b, int c)   // Stencils.h:13-57: face vertices ascending
{
    int q[3] = {a, b, c};
    std::sort(q, q + 3);
    return Key{{p, q[0], q[1], q[2]}};
}

Key ee_key(int a, int b, int c, int d)   // Stencils.h:86-117
{
    int p0 = std::min(a, b), p1 = std::max(a, b);
    int q0 = std::min(c, d), q1 = std::max(c, d);
    if (p0 > q0) {
        std::swap(p0, q0);
        std::swap(p1, q1);
    }
    return Key{{p0, p1, q0, q1}};
}

// ----------------------------------------------------------------------------------------------
// k-DOP broadphase: src/KDOPBroadPhase.{h,cpp}
// ----------------------------------------------------------------------------------------------
const int KAX = 13;
struct Axes {
    V3 a[KAX];
};
Axes make_axes()   // KDOPBroadPhase.cpp:10-32: 13 directions, each divided by its norm
{
    static const double raw[KAX][3] = {{1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},  {1, -1, 0},
                                       {0, 1, 1},  {0, 1, -1}, {1, 0, 1},  {1, 0, -1}, {1, 1, 1},
                                       {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};
    Axes ax;
    for (int i = 0; i < KAX; i++) {
        V3 v = load3(raw[i]);
        double n = norm(v);
        ax.a[i] = V3{v.x / n, v.y / n, v.z / n};
    }
    return ax;
}

struct Dop {
    double mn[KAX], mx[KAX];
};

// KDOPBroadPhase.cpp:43-74 (leaf bounds), generalised to `naxes` = 13 (k-DOP) or 3 (AABB,
// AABBBroadPhase.cpp:9-40, which subtracts eta from the coordinate itself).
void leaf_dops(const int32_t *faces, int64_t F, const Hist &h, double eta, bool aabb, std::vector<Dop> &out)
{
    Axes ax = make_axes();
    int na = aabb ? 3 : KAX;
    out.resize(F);
    for (int64_t f = 0; f < F; f++) {
        Dop &d = out[f];
        for (int k = 0; k < KAX; k++) {
            d.mn[k] = std::numeric_limits<double>::infinity();
            d.mx[k] = -std::numeric_limits<double>::infinity();
        }
        for (int j = 0; j < 3; j++) {
            int v = faces[3 * f + j];
            for (int64_t e = h.off[v]; e < h.off[v + 1]; e++) {
                V3 p = load3(h.hp + 3 * e);
                for (int k = 0; k < na; k++) {
                    double s = aabb ? (k == 0 ? p.x : (k == 1 ? p.y : p.z)) : dot(p, ax.a[k]);
                    d.mn[k] = smin(s - eta, d.mn[k]);
                    d.mx[k] = smax(s + eta, d.mx[k]);
                }
            }
        }
    }
}

inline bool dop_overlap(const Dop &a, const Dop &b, int na)   // KDOPBroadPhase.cpp:132-137
{
    for (int k = 0; k < na; k++)
        if (a.mx[k] < b.mn[k] || b.mx[k] < a.mn[k]) return false;
    return true;
}

inline bool neighboring(const int32_t *faces, int64_t f, int64_t g)   // Mesh.cpp:13-20
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (faces[3 * f + i] == faces[3 * g + j]) return true;
    return false;
}

struct CandSets {
    std::vector<Key> vf, ee;
};

// KDOPBroadPhase.cpp:150-184: the 6 VF and 9 EE stencils of an overlapping, non-neighbouring leaf
// pair, skipping those whose 4 vertices are all fixed.
template <class Sink>
void emit_pair(const int32_t *faces, int64_t f, int64_t g, const uint8_t *fixed, Sink &sink)
{
    const int32_t *F0 = faces + 3 * f;
    const int32_t *G0 = faces + 3 * g;
    auto fx = [&](int v) { return fixed && fixed[v]; };
    for (int i = 0; i < 3; i++) {
        bool lf = fx(F0[i]) && fx(G0[0]) && fx(G0[1]) && fx(G0[2]);
        bool rf = fx(G0[i]) && fx(F0[0]) && fx(F0[1]) && fx(F0[2]);
        if (!lf) sink.vf(vf_key(F0[i], G0[0], G0[1], G0[2]));
        if (!rf) sink.vf(vf_key(G0[i], F0[0], F0[1], F0[2]));
        for (int j = 0; j < 3; j++) {
            int a = F0[i], b = F0[(i + 1) % 3], c = G0[j], d = G0[(j + 1) % 3];
            bool ef = fx(a) && fx(b) && fx(c) && fx(d);
            if (!ef) sink.ee(ee_key(a, b, c, d));
        }
    }
}

struct VecSink {
    CandSets *cs;
    void vf(const Key &k) { cs->vf.push_back(k); }
    void ee(const Key &k) { cs->ee.push_back(k); }
};
struct SetSink {
    std::set<Key> *vfs, *ees;
    void vf(const Key &k) { vfs->insert(k); }
    void ee(const Key &k) { ees->insert(k); }
};

void sort_unique(std::vector<Key> &v)
{
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

// Brute force over all face pairs f < g.  The reference's tree visits exactly the leaf pairs whose
// bounds overlap (interior bounds are exact unions of child bounds, KDOPBroadPhase.cpp:85-98), in
// both orders, and emission is symmetric, so this enumerates the same stencil set.
void candidates_bruteforce(const int32_t *faces, int64_t F, const std::vector<Dop> &dops, int na,
                           const uint8_t *fixed, CandSets &out, std::vector<int64_t> *pairs)
{
    VecSink sink{&out};
    for (int64_t f = 0; f < F; f++)
        for (int64_t g = f + 1; g < F; g++) {
            if (!dop_overlap(dops[f], dops[g], na)) continue;
            if (neighboring(faces, f, g)) continue;
            if (pairs) {
                pairs->push_back(f);
                pairs->push_back(g);
            }
            emit_pair(faces, f, g, fixed, sink);
        }
    sort_unique(out.vf);
    sort_unique(out.ee);
}

// The reference-structured tree: top-down split of the node's children at the median of the
// longest-extent axis (KDOPBroadPhase.cpp:76-128; AABB variant AABBBroadPhase.cpp:42-89 with
// >= ties), recursive dual traversal (:130-149) and std::set insertion (:150-184).
struct Node {
    Dop d;
    int face;   // >= 0 for a leaf
    int left, right;
};

struct Tree {
    std::vector<Node> nodes;
    int na;
    int build(std::vector<int> &items)   // items: node indices (leaves) to group
    {
        int n = (int)items.size();
        if (n == 1) return items[0];
        Node node;
        node.face = -1;
        for (int k = 0; k < KAX; k++) {
            node.d.mn[k] = std::numeric_limits<double>::infinity();
            node.d.mx[k] = -std::numeric_limits<double>::infinity();
        }
        for (int it : items)
            for (int k = 0; k < na; k++) {
                node.d.mn[k] = smin(nodes[it].d.mn[k], node.d.mn[k]);
                node.d.mx[k] = smax(nodes[it].d.mx[k], node.d.mx[k]);
            }
        int axis = -1;
        if (na == KAX) {
            double best = 0;
            for (int k = 0; k < na; k++) {
                double len = node.d.mx[k] - node.d.mn[k];
                if (len > best) {
                    best = len;
                    axis = k;
                }
            }
            if (axis < 0) axis = 0;   // all extents <= 0 is unreachable for eta > 0
        } else {
            double l0 = node.d.mx[0] - node.d.mn[0], l1 = node.d.mx[1] - node.d.mn[1],
                   l2 = node.d.mx[2] - node.d.mn[2];
            if (l0 >= l1 && l0 >= l2)
                axis = 0;
            else if (l1 >= l0 && l1 >= l2)
                axis = 1;
            else
                axis = 2;
        }
        const std::vector<Node> &nd = nodes;
        std::sort(items.begin(), items.end(),
                  [&nd, axis](int a, int b) { return nd[a].d.mn[axis] < nd[b].d.mn[axis]; });
        std::vector<int> L(items.begin(), items.begin() + n / 2), R(items.begin() + n / 2, items.end());
        int l = build(L);
        int r = build(R);
        node.left = l;
        node.right = r;
        nodes.push_back(node);
        return (int)nodes.size() - 1;
    }
};

template <class Sink>
void dual_traverse(const Tree &T, int a, int b, const int32_t *faces, const uint8_t *fixed, Sink &sink)
{
    const Node &A = T.nodes[a];
    const Node &B = T.nodes[b];
    if (!dop_overlap(A.d, B.d, T.na)) return;
    if (A.face < 0) {
        dual_traverse(T, A.left, b, faces, fixed, sink);
        dual_traverse(T, A.right, b, faces, fixed, sink);
    } else if (B.face < 0) {
        dual_traverse(T, a, B.left, faces, fixed, sink);
        dual_traverse(T, a, B.right, faces, fixed, sink);
    } else {
        if (neighboring(faces, A.face, B.face)) return;
        emit_pair(faces, A.face, B.face, fixed, sink);
    }
}

// Worker threads of the all-core CPU baseline (orc_set_threads); 1 = the reference's single thread.
static int g_threads = 1;

// The dual traversal split into independent subtree pairs (the recursion's own branches, expanded
// breadth-first until there are plenty per thread), each traversed by one thread into its own vector,
// then merged and made unique: the same set as the single-threaded traversal.
void dual_traverse_parallel(const Tree &T, int root, const int32_t *faces, const uint8_t *fixed, CandSets &out)
{
    std::vector<std::pair<int, int>> tasks{{root, root}}, leaves;
    const size_t want = (size_t)64 * (size_t)g_threads;
    while (!tasks.empty() && tasks.size() + leaves.size() < want) {
        std::vector<std::pair<int, int>> next;
        for (auto [a, b] : tasks) {
            const Node &A = T.nodes[a], &B = T.nodes[b];
            if (!dop_overlap(A.d, B.d, T.na)) continue;
            if (A.face < 0) {
                next.push_back({A.left, b});
                next.push_back({A.right, b});
            } else if (B.face < 0) {
                next.push_back({a, B.left});
                next.push_back({a, B.right});
            } else {
                leaves.push_back({a, b});
            }
        }
        tasks.swap(next);
    }
    tasks.insert(tasks.end(), leaves.begin(), leaves.end());
    std::vector<CandSets> part(g_threads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)tasks.size(); i++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        VecSink sink{&part[tid]};
        dual_traverse(T, tasks[i].first, tasks[i].second, faces, fixed, sink);
    }
    for (auto &p : part) {
        out.vf.insert(out.vf.end(), p.vf.begin(), p.vf.end());
        out.ee.insert(out.ee.end(), p.ee.begin(), p.ee.end());
    }
    sort_unique(out.vf);
    sort_unique(out.ee);
}

void candidates_tree(const int32_t *faces, int64_t F, const std::vector<Dop> &dops, int na,
                     const uint8_t *fixed, CandSets &out, bool use_std_set, bool parallel = false)
{
    if (F == 0) return;
    Tree T;
    T.na = na;
    T.nodes.reserve(2 * F);
    std::vector<int> items(F);
    for (int64_t f = 0; f < F; f++) {
        Node n;
        n.d = dops[f];
        n.face = (int)f;
        n.left = n.right = -1;
        T.nodes.push_back(n);
        items[f] = (int)f;
    }
    int root = T.build(items);
    if (parallel) {
        dual_traverse_parallel(T, root, faces, fixed, out);
    } else if (use_std_set) {
        std::set<Key> vfs, ees;
        SetSink sink{&vfs, &ees};
        dual_traverse(T, root, root, faces, fixed, sink);
        out.vf.assign(vfs.begin(), vfs.end());
        out.ee.assign(ees.begin(), ees.end());
    } else {
        VecSink sink{&out};
        dual_traverse(T, root, root, faces, fixed, sink);
        sort_unique(out.vf);
        sort_unique(out.ee);
    }
}

// ----------------------------------------------------------------------------------------------
// Proximity: include/Distance.h:14-174 and src/Distance.cpp:12-66
// ----------------------------------------------------------------------------------------------
bool vertex_plane_lt(V3 p, V3 q0, V3 q1, V3 q2, double eta)   // Distance.h:14-19
{
    V3 c = cross(q1 - q0, q2 - q0);
    return dot(c, p - q0) * dot(c, p - q0) < eta * eta * dot(c, c);
}
bool line_line_lt(V3 p0, V3 p1, V3 q0, V3 q1, double eta)   // Distance.h:23-28
{
    V3 c = cross(p1 - p0, q1 - q0);
    return dot(c, q0 - p0) * dot(c, q0 - p0) < eta * eta * dot(c, c);
}
double clamp_unit(double u) { return smin(1.0, smax(u, 0.0)); }   // Distance.h:183-186

V3 vertex_face_distance(V3 p, V3 q0, V3 q1, V3 q2, double *bary)   // Distance.h:32-115
{
    V3 ab = q1 - q0, ac = q2 - q0, ap = p - q0;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) {
        bary[0] = 1.0; bary[1] = 0.0; bary[2] = 0.0;
        return q0 - p;
    }
    V3 bp = p - q1;
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
        bary[0] = 0.0; bary[1] = 1.0; bary[2] = 0.0;
        return q1 - p;
    }
    double vc = d1 * d4 - d3 * d2;
    if ((vc <= 0) && (d1 >= 0) && (d3 <= 0)) {
        double v = d1 / (d1 - d3);
        bary[0] = 1.0 - v; bary[1] = v; bary[2] = 0;
        return (q0 + v * ab) - p;
    }
    V3 cp = p - q2;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) {
        bary[0] = 0; bary[1] = 0; bary[2] = 1.0;
        return q2 - p;
    }
    double vb = d5 * d2 - d1 * d6;
    if ((vb <= 0) && (d2 >= 0) && (d6 <= 0)) {
        double w = d2 / (d2 - d6);
        bary[0] = 1 - w; bary[1] = 0; bary[2] = w;
        return (q0 + w * ac) - p;
    }
    double va = d3 * d6 - d5 * d4;
    if ((va <= 0) && (d4 - d3 >= 0) && (d5 - d6 >= 0)) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary[0] = 0; bary[1] = 1.0 - w; bary[2] = w;
        return (q1 + w * (q2 - q1)) - p;
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    double u = 1.0 - v - w;
    bary[0] = u; bary[1] = v; bary[2] = w;
    return ((u * q0 + v * q1) + w * q2) - p;
}

V3 edge_edge_distance(V3 p0, V3 p1, V3 q0, V3 q1, double *bary)   // Distance.h:119-174
{
    V3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double s, t;
    double c = dot(d1, r), b = dot(d1, d2);
    double denom = a * e - b * b;
    if (denom != 0.0)
        s = clamp_unit((b * f - c * e) / denom);
    else
        s = 0;
    double tnom = b * s + f;
    if (tnom < 0 || e == 0) {
        t = 0;
        if (a == 0) s = 0; else s = clamp_unit(-c / a);
    } else if (tnom > e) {
        t = 1.0;
        if (a == 0) s = 0; else s = clamp_unit((b - c) / a);
    } else
        t = tnom / e;
    V3 c1 = p0 + s * d1, c2 = q0 + t * d2;
    bary[0] = 1.0 - s; bary[1] = s; bary[2] = 1.0 - t; bary[3] = t;
    return c2 - c1;
}

// Distance.cpp:12-66
double mesh_self_distance(const double *verts, int64_t V, const int32_t *faces, int64_t F, const uint8_t *fixed,
                          int64_t *nvf, int64_t *nee)
{
    double closest = std::numeric_limits<double>::infinity();
    for (int64_t i = 0; i < V; i++)
        for (int64_t j = 0; j < F; j++) {
            const int32_t *fj = faces + 3 * j;
            if (fj[0] == i || fj[1] == i || fj[2] == i) continue;   // Mesh::vertexOfFace
            for (int k = 0; k < 3; k++) {
                V3 d = load3(verts + 3 * i) - load3(verts + 3 * (int64_t)fj[k]);
                double dist = dot(d, d);
                if (dist < closest) closest = dist;
            }
        }
    closest = std::sqrt(closest);
    // History(verts) + finishHistory(verts): two identical entries per vertex.
    std::vector<int64_t> off(V + 1);
    std::vector<double> ht(2 * V), hp(6 * V);
    for (int64_t v = 0; v <= V; v++) off[v] = 2 * v;
    for (int64_t v = 0; v < V; v++) {
        ht[2 * v] = 0.0;
        ht[2 * v + 1] = 1.0;
        for (int k = 0; k < 3; k++) hp[6 * v + k] = hp[6 * v + 3 + k] = verts[3 * v + k];
    }
    Hist h{off.data(), ht.data(), hp.data()};
    std::vector<Dop> d;
    leaf_dops(faces, F, h, closest, true, d);
    CandSets cs;
    candidates_bruteforce(faces, F, d, 3, fixed, cs, nullptr);
    *nvf = (int64_t)cs.vf.size();
    *nee = (int64_t)cs.ee.size();
    closest = std::numeric_limits<double>::infinity();
    double bary[4];
    for (const Key &k : cs.vf) {
        double dist = norm(vertex_face_distance(load3(verts + 3 * (int64_t)k.k[0]), load3(verts + 3 * (int64_t)k.k[1]),
                                                load3(verts + 3 * (int64_t)k.k[2]), load3(verts + 3 * (int64_t)k.k[3]), bary));
        if (dist < closest) closest = dist;
    }
    for (const Key &k : cs.ee) {
        double dist = norm(edge_edge_distance(load3(verts + 3 * (int64_t)k.k[0]), load3(verts + 3 * (int64_t)k.k[1]),
                                              load3(verts + 3 * (int64_t)k.k[2]), load3(verts + 3 * (int64_t)k.k[3]), bary));
        if (dist < closest) closest = dist;
    }
    return closest;
}

// Result holder handed across the C boundary.
// ----------------------------------------------------------------------------------------------
// SeparatingPlaneNarrowPhase: src/SeparatingPlaneNarrowPhase.cpp:9-278, History.cpp:42-93.
// Written in the reference's own shape (recursive bisection, the History lookups as the reference
// performs them); the device version is an explicit-stack restatement, which these tests check.
// ----------------------------------------------------------------------------------------------
// History::getPosAtTime (History.cpp:42-76) on vertex v: idx is relative to the vertex's first entry.
void get_pos_at_time(const Hist &h, int v, double time, V3 &pos, int &idx, int64_t start, int64_t end)
{
    const int64_t b = h.off[v], size = h.off[v + 1] - b;
    if (end - start < 10) {
        int64_t next = start;
        while (next < size && h.ht[b + next] <= time) next++;
        int64_t prev = next - 1;
        if (next == size) {
            pos = load3(h.hp + 3 * (b + prev));
            idx = (int)prev;
            return;
        }
        double dt = h.ht[b + next] - h.ht[b + prev];
        double alpha = (time - h.ht[b + prev]) / dt;
        pos = (1.0 - alpha) * load3(h.hp + 3 * (b + prev)) + alpha * load3(h.hp + 3 * (b + next));
        idx = (int)prev;
        return;
    }
    int64_t mid = (start + end) / 2;
    if (h.ht[b + mid] < time)
        return get_pos_at_time(h, v, time, pos, idx, mid, end);
    return get_pos_at_time(h, v, time, pos, idx, start, mid);
}
void get_pos_at_time(const Hist &h, int v, double time, V3 &pos, int &idx)
{
    get_pos_at_time(h, v, time, pos, idx, 0, h.off[v + 1] - h.off[v] - 1);
}

// History::computeMinimumGap (History.cpp:78-93)
double compute_minimum_gap(const Hist &h, int64_t nverts)
{
    double result = 1.0;
    for (int64_t i = 0; i < nverts; i++)
        for (int64_t j = h.off[i] + 1; j < h.off[i + 1]; j++) result = std::min(h.ht[j] - h.ht[j - 1], result);
    return result;
}

double plane_intersect(V3 planePos, V3 planeVel, V3 planeNormal, V3 ptOld, V3 ptNew, double ptdt, double eta)
{   // :267-278
    double numerator = 0.5 * eta - dot(ptOld - planePos, planeNormal);
    double denom = -dot(planeVel, planeNormal);
    if (ptdt != 0.0) denom += dot(ptNew - ptOld, planeNormal) / ptdt;
    return numerator / denom;
}

double plane_trajectory_intersect(const Hist &h, int v, int startidx, bool forward, V3 planePos, V3 planeVel,
                                  V3 planeNormal, V3 ptstart, double timestart, double eta)
{   // :216-265
    const int64_t b = h.off[v];
    const int size = (int)(h.off[v + 1] - b);
    auto T = [&](int i) { return h.ht[b + i]; };
    auto P = [&](int i) { return load3(h.hp + 3 * (b + i)); };
    double dt = (forward ? T(startidx) - timestart : timestart - T(startidx));
    V3 newpos = P(startidx);
    double t = plane_intersect(planePos, (forward ? 1.0 : -1.0) * planeVel, planeNormal, ptstart, newpos, dt, eta);
    if (t >= 0 && t <= dt) return t;
    if (forward) {
        for (int i = startidx; i < size; i++) {
            if (i == size - 1) return std::numeric_limits<double>::infinity();
            int next = i + 1;
            V3 newplanepos = planePos + (T(i) - timestart) * planeVel;
            double d = T(next) - T(i);
            double u = plane_intersect(newplanepos, planeVel, planeNormal, P(i), P(next), d, eta);
            if (u >= 0 && u <= d) return T(i) + u - timestart;
        }
    } else {
        for (int i = startidx; i >= 0; i--) {
            if (i == 0) return std::numeric_limits<double>::infinity();
            int prev = i - 1;
            V3 newplanepos = planePos - (timestart - T(i)) * planeVel;
            double d = T(i) - T(prev);
            double u = plane_intersect(newplanepos, -1.0 * planeVel, planeNormal, P(i), P(prev), d, eta);
            if (u >= 0 && u <= d) return timestart - (T(i) - u);
        }
    }
    return std::numeric_limits<double>::infinity();
}

bool sp_leaf(bool vf, const V3 *o, const V3 *n, double eta, Ctx *ctx)
{   // :76-140: the primitive tests in the reference's order on one window
    double t;
    if (vf) {
        if (vertex_face_ctcd(o[0], o[1], o[2], o[3], n[0], n[1], n[2], n[3], eta, t, ctx)) return true;
        for (int edge = 0; edge < 3; edge++) {
            int a = 1 + (edge % 3), c = 1 + ((edge + 1) % 3);
            if (vertex_edge_ctcd(o[0], o[a], o[c], n[0], n[a], n[c], eta, t, ctx)) return true;
        }
        for (int vert = 0; vert < 3; vert++)
            if (vertex_vertex_ctcd(o[0], o[1 + vert], n[0], n[1 + vert], eta, t, ctx)) return true;
        return false;
    }
    if (edge_edge_ctcd(o[0], o[1], o[2], o[3], n[0], n[1], n[2], n[3], eta, t, ctx)) return true;
    static const int ve_idx[4][3] = {{0, 2, 3}, {1, 2, 3}, {2, 0, 1}, {3, 0, 1}};
    static const int vv_idx[4][2] = {{0, 2}, {0, 3}, {1, 2}, {1, 3}};
    for (int k = 0; k < 4; k++) {
        const int *q = ve_idx[k];
        if (vertex_edge_ctcd(o[q[0]], o[q[1]], o[q[2]], n[q[0]], n[q[1]], n[q[2]], eta, t, ctx)) return true;
    }
    for (int k = 0; k < 4; k++) {
        const int *q = vv_idx[k];
        if (vertex_vertex_ctcd(o[q[0]], o[q[1]], n[q[0]], n[q[1]], eta, t, ctx)) return true;
    }
    return false;
}

// A call whose interval equals its caller's never returns in the reference (checkInterval is a pure
// function of the interval): such a stencil is reported undecided (*undecided = 1) rather than
// crashing the process as the reference would.  So is one whose recursion goes deeper than
// SP_MAX_DEPTH levels (the product's bound; each level at most halves the interval, so only an
// interval that closes in on a history entry time gets that deep).
constexpr int SP_MAX_DEPTH = 160;
bool sp_check_interval(bool vf, const Hist &h, const int *verts, double eta, double mint, double maxt, double eps,
                       Ctx *ctx, int depth, int &maxdepth, int *undecided)
{   // :63-214
    if (depth > maxdepth) maxdepth = depth;
    if (depth > SP_MAX_DEPTH) {
        *undecided = 1;
        return false;
    }
    if (maxt < mint) return false;
    if (maxt - mint < 3 * eps) {
        bool ok = true;
        V3 oldpos[4], newpos[4];
        for (int i = 0; i < 4; i++) {
            int oldidx, newidx;
            get_pos_at_time(h, verts[i], mint, oldpos[i], oldidx);
            get_pos_at_time(h, verts[i], maxt, newpos[i], newidx);
            if (oldidx != newidx) {
                ok = false;
                break;
            }
        }
        if (ok) return sp_leaf(vf, oldpos, newpos, eta, ctx);
    }
    double midt = 0.5 * (mint + maxt);
    V3 midpos[4];
    int mididx[4];
    for (int i = 0; i < 4; i++) get_pos_at_time(h, verts[i], midt, midpos[i], mididx[i]);
    double bary[4] = {0, 0, 0, 0};
    V3 closest = vf ? vertex_face_distance(midpos[0], midpos[1], midpos[2], midpos[3], bary)
                    : edge_edge_distance(midpos[0], midpos[1], midpos[2], midpos[3], bary);
    double distsq = dot(closest, closest);
    if (distsq < eta * eta) return true;
    V3 prevpos[4], nextpos[4];
    double dts[4];
    for (int i = 0; i < 4; i++) {
        const int64_t k = h.off[verts[i]] + mididx[i];
        prevpos[i] = load3(h.hp + 3 * k);
        nextpos[i] = load3(h.hp + 3 * (k + 1));
        dts[i] = h.ht[k + 1] - h.ht[k];
    }
    auto div = [](V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; };
    V3 planepos, planevel;
    if (vf) {
        planepos = 0.5 * (((midpos[0] + bary[0] * midpos[1]) + bary[1] * midpos[2]) + bary[2] * midpos[3]);
        planevel = 0.5 * (((div(nextpos[0] - prevpos[0], dts[0]) + div(bary[0] * (nextpos[1] - prevpos[1]), dts[1])) +
                           div(bary[1] * (nextpos[2] - prevpos[2]), dts[2])) +
                          div(bary[2] * (nextpos[3] - prevpos[3]), dts[3]));
    } else {
        planepos = 0.5 * (((bary[0] * midpos[0] + bary[1] * midpos[1]) + bary[2] * midpos[2]) + bary[3] * midpos[3]);
        planevel = 0.5 * (((div(bary[0] * (nextpos[0] - prevpos[0]), dts[0]) + div(bary[1] * (nextpos[1] - prevpos[1]), dts[1])) +
                           div(bary[2] * (nextpos[2] - prevpos[2]), dts[2])) +
                          div(bary[3] * (nextpos[3] - prevpos[3]), dts[3]));
    }
    closest = div(closest, norm(closest));
    double t = std::numeric_limits<double>::infinity();
    for (int vert = 0; vert < 4; vert++) {
        double sign = (vert == 0 || (vert == 1 && !vf)) ? -1.0 : 1.0;
        t = std::min(t, plane_trajectory_intersect(h, verts[vert], mididx[vert], false, planepos, planevel,
                                                   sign * closest, midpos[vert], midt, eta));
    }
    if (t < 1.0) {
        double newmaxt = midt - std::max(0.0, t - eps);
        if (newmaxt == maxt) {
            *undecided = 1;
            return false;
        }
        if (sp_check_interval(vf, h, verts, eta, mint, newmaxt, eps, ctx, depth + 1, maxdepth, undecided)) return true;
        if (*undecided) return false;
    }
    t = std::numeric_limits<double>::infinity();
    for (int vert = 0; vert < 4; vert++) {
        double sign = (vert == 0 || (vert == 1 && !vf)) ? -1.0 : 1.0;
        t = std::min(t, plane_trajectory_intersect(h, verts[vert], mididx[vert] + 1, true, planepos, planevel,
                                                   sign * closest, midpos[vert], midt, eta));
    }
    if (t < 1.0) {
        double newmint = midt + std::max(0.0, t - eps);
        if (newmint == mint) {
            *undecided = 1;
            return false;
        }
        if (sp_check_interval(vf, h, verts, eta, newmint, maxt, eps, ctx, depth + 1, maxdepth, undecided)) return true;
        if (*undecided) return false;
    }
    return false;
}

struct Result {
    CandSets cs;
    std::vector<int64_t> pairs;
};

// ----------------------------------------------------------------------------------------------
// Velocity filter: include/VelocityFilter.h:7-45, src/VelocityFilter.cpp:16-87, src/ActiveLayers.cpp,
// src/PenaltyGroup.cpp, src/PenaltyPotential.cpp, src/SimulationState.h.  Written in the reference's
// own shape: full-length state and force vectors per firing, std::set group stencils, std::map layer
// depths and a std::priority_queue of groups.  The collision queries are this file's brute-force k-DOP
// broadphase and recursive separating-plane narrowphase.
// ----------------------------------------------------------------------------------------------
struct VHist {   // src/History.h:7-42: per-vertex (time, pos) entries
    std::vector<std::vector<std::pair<double, V3>>> e;
    explicit VHist(const std::vector<double> &q) : e(q.size() / 3)
    {   // History.cpp:8-16
        for (size_t v = 0; v < e.size(); v++) e[v].push_back({0.0, load3(&q[3 * v])});
    }
    void finish(const std::vector<double> &q)
    {   // History.cpp:31-39
        for (size_t v = 0; v < e.size(); v++) e[v].push_back({1.0, load3(&q[3 * v])});
    }
    int64_t entries() const
    {
        int64_t n = 0;
        for (const auto &x : e) n += (int64_t)x.size();
        return n;
    }
    void csr(std::vector<int64_t> &off, std::vector<double> &t, std::vector<double> &p) const
    {
        off.assign(e.size() + 1, 0);
        t.clear();
        p.clear();
        for (size_t v = 0; v < e.size(); v++) {
            for (const auto &x : e[v]) {
                t.push_back(x.first);
                p.push_back(x.second.x);
                p.push_back(x.second.y);
                p.push_back(x.second.z);
            }
            off[v + 1] = (int64_t)t.size();
        }
    }
};

struct PGroup {   // PenaltyGroup.h:12-42, PenaltyGroup.cpp:7-71
    std::vector<Key> vfs, ees;
    std::set<int> verts;
    int nextstep = 1;
    double dt, outer, inner, k, cor;
    double next_fire() const { return nextstep * dt; }
};
struct PGroupCmp {   // ActiveLayers.cpp:18-21
    bool operator()(const PGroup *a, const PGroup *b) const { return b->next_fire() < a->next_fire(); }
};

inline void add_seg(std::vector<double> &F, int v, double s, V3 f)
{   // F.segment<3>(3v) += s * f
    F[3 * v] = F[3 * v] + s * f.x;
    F[3 * v + 1] = F[3 * v + 1] + s * f.y;
    F[3 * v + 2] = F[3 * v + 2] + s * f.z;
}
inline void sub_seg(std::vector<double> &F, int v, double s, V3 f)
{   // F.segment<3>(3v) -= s * f  (s = 1 for the unscaled form, exact)
    F[3 * v] = F[3 * v] - s * f.x;
    F[3 * v + 1] = F[3 * v + 1] - s * f.y;
    F[3 * v + 2] = F[3 * v + 2] - s * f.z;
}

// VertexFacePenaltyPotential::addForce (PenaltyPotential.cpp:9-31)
bool vf_penalty(const std::vector<double> &q, const std::vector<double> &v, std::vector<double> &F, const Key &s,
                double outer, double inner, double k, double cor)
{
    V3 p = load3(&q[3 * (int64_t)s.k[0]]), a = load3(&q[3 * (int64_t)s.k[1]]), b = load3(&q[3 * (int64_t)s.k[2]]),
       c = load3(&q[3 * (int64_t)s.k[3]]);
    if (!vertex_plane_lt(p, a, b, c, outer)) return false;
    double bary[4];
    V3 cv = vertex_face_distance(p, a, b, c, bary);
    double dist = norm(cv);
    if (dist >= outer || dist < inner) return false;
    V3 vp = load3(&v[3 * (int64_t)s.k[0]]), va = load3(&v[3 * (int64_t)s.k[1]]), vb = load3(&v[3 * (int64_t)s.k[2]]),
       vc = load3(&v[3 * (int64_t)s.k[3]]);
    V3 rel = ((V3{-vp.x, -vp.y, -vp.z} + bary[0] * va) + bary[1] * vb) + bary[2] * vc;
    if (dot(rel, cv) > 0) k *= cor;
    double sc = k * (outer - dist) / (outer - inner);
    V3 lf{(sc * cv.x) / dist, (sc * cv.y) / dist, (sc * cv.z) / dist};
    F[3 * (int64_t)s.k[0]] -= lf.x;
    F[3 * (int64_t)s.k[0] + 1] -= lf.y;
    F[3 * (int64_t)s.k[0] + 2] -= lf.z;
    add_seg(F, s.k[1], bary[0], lf);
    add_seg(F, s.k[2], bary[1], lf);
    add_seg(F, s.k[3], bary[2], lf);
    return true;
}

// EdgeEdgePenaltyPotential::addForce (PenaltyPotential.cpp:34-61)
bool ee_penalty(const std::vector<double> &q, const std::vector<double> &v, std::vector<double> &F, const Key &s,
                double outer, double inner, double k, double cor)
{
    V3 p0 = load3(&q[3 * (int64_t)s.k[0]]), p1 = load3(&q[3 * (int64_t)s.k[1]]), q0 = load3(&q[3 * (int64_t)s.k[2]]),
       q1 = load3(&q[3 * (int64_t)s.k[3]]);
    if (!line_line_lt(p0, p1, q0, q1, outer)) return false;
    double bary[4];
    V3 cv = edge_edge_distance(p0, p1, q0, q1, bary);
    double dist = norm(cv);
    if (dist >= outer || dist < inner) return false;
    V3 v0 = load3(&v[3 * (int64_t)s.k[0]]), v1 = load3(&v[3 * (int64_t)s.k[1]]), v2 = load3(&v[3 * (int64_t)s.k[2]]),
       v3_ = load3(&v[3 * (int64_t)s.k[3]]);
    V3 rel = (((-bary[0]) * v0 - bary[1] * v1) + bary[2] * v2) + bary[3] * v3_;
    if (dot(rel, cv) > 0) k *= cor;
    double sc = k * (outer - dist) / (outer - inner);
    V3 lf{(sc * cv.x) / dist, (sc * cv.y) / dist, (sc * cv.z) / dist};
    sub_seg(F, s.k[0], bary[0], lf);
    sub_seg(F, s.k[1], bary[1], lf);
    add_seg(F, s.k[2], bary[2], lf);
    add_seg(F, s.k[3], bary[3], lf);
    return true;
}

struct SimState {   // SimulationState.h
    std::vector<double> q, v, minv, last;
};

struct VFStats {
    int64_t iterations = 0, layers = 0, firings = 0, history_entries = 0, vf_candidates = 0, ee_candidates = 0,
            vf_collisions = 0, ee_collisions = 0, undecided = 0;
    double closest = 0;
};

class Layers {   // ActiveLayers.h:23-75, ActiveLayers.cpp:23-256
public:
    Layers(double outer, double inner, double dt, double k, double term, double cor, const int32_t *faces,
           int64_t F)
        : outer_(outer), inner_(inner), dt_(dt), k_(k), term_(term), cor_(cor), faces_(faces), F_(F)
    {
    }
    ~Layers()
    {
        for (PGroup *g : groups_) delete g;
    }
    // ActiveLayers.cpp:217-256.  Returns 1 when no collision was found, 0 otherwise, -1 when the
    // separating-plane recursion does not terminate for some stencil (the reference overflows its stack).
    int run_one_iteration(SimState &s, VFStats &st, VHist **final_hist)
    {
        delete hist_;
        hist_ = new VHist(s.q);
        while (!step(s, st)) {
        }
        st.history_entries = hist_->entries();
        std::set<int> fixed;
        for (size_t i = 0; i < s.minv.size(); i++)
            if (s.minv[i] == 0.0) fixed.insert((int)(i / 3));
        std::vector<Key> vfs, ees;
        int rc = collision_detection(fixed, vfs, ees, st);
        if (rc < 0) return -1;
        rollback();
        for (const Key &k : vfs) add_stencil(k, false);
        for (const Key &k : ees) add_stencil(k, true);
        if (final_hist) *final_hist = hist_;
        return (vfs.empty() && ees.empty()) ? 1 : 0;
    }
    int64_t layers() const { return deepest_; }

private:
    double layer_depth(int layer) const { return inner_ + (outer_ - inner_) / double(layer); }   // :167-170

    void add_groups(int maxdepth)
    {   // :67-83
        while (deepest_ < maxdepth) {
            int depth = deepest_ + 1;
            double fudge = 1e-4;
            PGroup *g = new PGroup;
            g->dt = dt_ / double(depth) / std::sqrt(double(depth) + fudge);
            g->outer = layer_depth(depth);
            g->inner = inner_;
            g->k = k_ * depth * depth * depth;
            g->cor = cor_;
            groups_.push_back(g);
            queue_.push(g);
            deepest_++;
        }
    }

    void add_stencil(const Key &k, bool ee)
    {   // :41-65
        int &d = (ee ? eedepth_ : vfdepth_)[k];
        int olddepth = d;
        for (int i = 0; i < 5; i++) {
            add_groups(olddepth + i + 1);
            PGroup *g = groups_[olddepth + i];
            (ee ? g->ees : g->vfs).push_back(k);
            for (int j = 0; j < 4; j++) g->verts.insert(k.k[j]);
            d++;
        }
    }

    bool step(SimState &s, VFStats &st)
    {   // :85-150
        size_t n = s.q.size();
        std::vector<double> F(n), newq(n, 0.0), newv(n, 0.0);
        if (!queue_.empty()) {
            PGroup *g = queue_.top();
            double newtime = g->next_fire();
            if (newtime < term_) {
                queue_.pop();
                st.firings++;
                for (int vert : g->verts)
                    for (int j = 0; j < 3; j++) {
                        size_t i = 3 * (size_t)vert + j;
                        double dt = newtime - s.last[i];
                        newq[i] = s.q[i] + (newtime - s.last[i]) * s.v[i];
                        newv[i] = (newq[i] - s.q[i]) / dt;
                    }
                std::fill(F.begin(), F.end(), 0.0);
                std::vector<double> gf(n, 0.0);   // PenaltyGroup::addForce (PenaltyGroup.cpp:35-53)
                for (const Key &k : g->vfs) vf_penalty(newq, newv, gf, k, g->outer, g->inner, g->k, g->cor);
                for (const Key &k : g->ees) ee_penalty(newq, newv, gf, k, g->outer, g->inner, g->k, g->cor);
                for (size_t i = 0; i < n; i++) F[i] = F[i] + gf[i] * g->dt;
                for (int vert : g->verts) {
                    bool touched = false;
                    for (int j = 0; j < 3; j++)
                        if (F[3 * (size_t)vert + j] != 0.0) touched = true;
                    if (!touched) continue;
                    for (int j = 0; j < 3; j++) {
                        size_t i = 3 * (size_t)vert + j;
                        s.v[i] = s.v[i] + s.minv[i] * F[i];
                        s.q[i] = newq[i];
                        s.last[i] = newtime;
                    }
                    hist_->e[vert].push_back({newtime, load3(&s.q[3 * (size_t)vert])});
                }
                g->nextstep++;
                queue_.push(g);
                return false;
            }
        }
        for (size_t i = 0; i < n; i++) {
            s.q[i] = s.q[i] + (term_ - s.last[i]) * s.v[i];
            s.last[i] = term_;
        }
        hist_->finish(s.q);
        return true;
    }

    void rollback()
    {   // :152-165
        for (PGroup *g : groups_) g->nextstep = 1;
        std::priority_queue<PGroup *, std::vector<PGroup *>, PGroupCmp> nq;
        while (!queue_.empty()) {
            nq.push(queue_.top());
            queue_.pop();
        }
        queue_ = nq;
    }

    double thickness(const Key &k, bool ee) const
    {   // :172-186
        const std::map<Key, int> &m = ee ? eedepth_ : vfdepth_;
        auto it = m.find(k);
        return it != m.end() ? layer_depth(it->second + 1) : layer_depth(1);
    }

    int collision_detection(const std::set<int> &fixedset, std::vector<Key> &vfs, std::vector<Key> &ees,
                            VFStats &st)
    {   // :188-215: KDOPBroadPhase over the history, then SeparatingPlaneNarrowPhase with per-stencil eta
        std::vector<int64_t> off;
        std::vector<double> ht, hp;
        hist_->csr(off, ht, hp);
        Hist h{off.data(), ht.data(), hp.data()};
        int64_t V = (int64_t)hist_->e.size();
        std::vector<uint8_t> fixed(V, 0);
        for (int v : fixedset) fixed[v] = 1;
        std::vector<Dop> d;
        leaf_dops(faces_, F_, h, outer_, false, d);
        CandSets cs;
        candidates_bruteforce(faces_, F_, d, KAX, fixedset.empty() ? nullptr : fixed.data(), cs, nullptr);
        st.vf_candidates += (int64_t)cs.vf.size();
        st.ee_candidates += (int64_t)cs.ee.size();
        double eps = compute_minimum_gap(h, V) / 4.0;   // SeparatingPlaneNarrowPhase.cpp:9-24
        int und = 0;
        for (int pass = 0; pass < 2; pass++) {
            bool ee = pass == 1;
            for (const Key &k : ee ? cs.ee : cs.vf) {
                Ctx ctx;
                int md = 0, u = 0;
                int verts[4] = {k.k[0], k.k[1], k.k[2], k.k[3]};
                if (sp_check_interval(!ee, h, verts, thickness(k, ee), 0.0, 1.0, eps, &ctx, 0, md, &u))
                    (ee ? ees : vfs).push_back(k);
                und += u;
            }
        }
        st.vf_collisions += (int64_t)vfs.size();
        st.ee_collisions += (int64_t)ees.size();
        st.undecided += und;
        return und ? -1 : 0;
    }

    double outer_, inner_, dt_, k_, term_, cor_;
    const int32_t *faces_;
    int64_t F_;
    int deepest_ = 0;
    std::vector<PGroup *> groups_;
    std::priority_queue<PGroup *, std::vector<PGroup *>, PGroupCmp> queue_;
    std::map<Key, int> vfdepth_, eedepth_;
    VHist *hist_ = nullptr;

public:
    Layers(const Layers &) = delete;
    Layers &operator=(const Layers &) = delete;
    void release_history(VHist **keep)
    {
        *keep = hist_;
        hist_ = nullptr;
    }
};

// VelocityFilter::velocityFilter (VelocityFilter.cpp:16-87).  Returns the reference's value (iterations,
// or -1 / -2 / -3), or -100 when the separating-plane recursion does not terminate for some stencil.
int velocity_filter(const double *qstart, double *qend, int64_t V, const int32_t *faces, int64_t F,
                    const double *invm, double outer, double inner, double stiff, double substep, double cor,
                    int max_rollbacks, VFStats &st, VHist **final_hist)
{
    for (int64_t i = 0; i < V; i++) {   // :24-37
        bool found = false;
        for (int64_t j = 0; j < F; j++)
            for (int k = 0; k < 3; k++)
                if (faces[3 * j + k] == i) found = true;
        if (!found) return -1;
    }
    double maxmass = 0;
    for (int64_t i = 0; i < 3 * V; i++) maxmass = smax(maxmass, invm[i]);
    double dt = substep;
    double k = stiff / dt * std::sqrt(maxmass);
    Layers al(outer, inner, dt, k, 1.0, cor, faces, F);
    std::vector<uint8_t> fixed(V, 0);
    bool any_fixed = false;
    for (int64_t i = 0; i < 3 * V; i++)
        if (invm[i] == 0.0) fixed[i / 3] = 1, any_fixed = true;
    int64_t nvf = 0, nee = 0;
    double closest = mesh_self_distance(qstart, V, faces, F, any_fixed ? fixed.data() : nullptr, &nvf, &nee);
    st.closest = closest;
    if (closest < inner) return -2;
    for (int iter = 0; iter < max_rollbacks; iter++) {
        SimState s;
        s.q.assign(qstart, qstart + 3 * V);
        s.v.resize(3 * V);
        for (int64_t i = 0; i < 3 * V; i++) s.v[i] = qend[i] - qstart[i];
        s.minv.assign(invm, invm + 3 * V);
        s.last.assign(3 * V, 0.0);
        st.iterations = iter + 1;
        int r = al.run_one_iteration(s, st, nullptr);
        st.layers = al.layers();
        if (r < 0) return -100;
        if (r == 1) {
            for (int64_t i = 0; i < 3 * V; i++) qend[i] = s.q[i];
            if (final_hist) al.release_history(final_hist);
            return iter;
        }
    }
    return -3;
}

}   // namespace

// ==============================================================================================
// C interface (ctypes).  Positions are flat double[3V]; faces flat int32[3F] (a column of
// Eigen::Matrix3Xi per face); histories CSR (off[V+1] int64, times[E], pos[3E]).
// ==============================================================================================
extern "C" {

void orc_set_dot_order(int order) { g_dot_order = order; }
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
void orc_set_rootfinder(void *fn) { g_rootfinder = (rootfinder_fn)fn; }

int orc_aberth_roots(const double *op, int degree, double *zr, double *zi)
{
    return aberth_roots(op, degree, zr, zi);
}

int orc_find_roots(double *op, int degree, double *zr, double *zi) { return find_roots(op, degree, zr, zi); }

// stitch() (History.cpp:98-140) for the nv <= 4 vertices verts[]: writes up to cap stitched times and
// nv positions per time (xyz, vertex-major within a time); returns the number of stitched times.
int64_t orc_stitch(const int64_t *off, const double *ht, const double *hp, const int32_t *verts, int nv, int64_t cap,
                   double *times, double *pos)
{
    Hist h{off, ht, hp};
    int v[4] = {0, 0, 0, 0};
    for (int i = 0; i < nv && i < 4; i++) v[i] = verts[i];
    std::vector<double> T;
    std::vector<V3> P;
    stitch(h, v, nv, T, P);
    const int64_t n = (int64_t)T.size();
    for (int64_t k = 0; k < n && k < cap; k++) {
        times[k] = T[k];
        for (int i = 0; i < nv; i++) {
            const V3 &q = P[k * nv + i];
            pos[3 * (k * nv + i)] = q.x;
            pos[3 * (k * nv + i) + 1] = q.y;
            pos[3 * (k * nv + i) + 2] = q.z;
        }
    }
    return n;
}

void orc_kdop_axes(double *out39)
{
    Axes ax = make_axes();
    for (int i = 0; i < KAX; i++) {
        out39[3 * i] = ax.a[i].x;
        out39[3 * i + 1] = ax.a[i].y;
        out39[3 * i + 2] = ax.a[i].z;
    }
}

int orc_get_quad_roots(double a, double b, double c, double *t0, double *t1)
{
    return get_quad_roots(a, b, c, *t0, *t1);
}

// findIntervals on a copy of op; writes up to cap intervals as (l,u) pairs; returns the count.
int orc_find_intervals(const double *op, int n, int pos, double *out_lu, int cap, double *out_op,
                       double *margin)
{
    double c[8];
    for (int i = 0; i <= n; i++) c[i] = op[i];
    std::vector<TI> iv;
    Ctx ctx;
    find_intervals(c, n, iv, pos != 0, &ctx);
    int m = (int)iv.size();
    for (int i = 0; i < m && i < cap; i++) {
        out_lu[2 * i] = iv[i].l;
        out_lu[2 * i + 1] = iv[i].u;
    }
    if (out_op)
        for (int i = 0; i <= n; i++) out_op[i] = c[i];
    if (margin) *margin = ctx.margin;
    return m;
}

// The four primitives on 3-vectors packed as start positions then end positions.
int orc_vertex_face_ctcd(const double *q, double eta, double *t, double *margin)
{
    Ctx ctx;
    double tt = 0;
    bool r = vertex_face_ctcd(load3(q), load3(q + 3), load3(q + 6), load3(q + 9), load3(q + 12),
                              load3(q + 15), load3(q + 18), load3(q + 21), eta, tt, &ctx);
    if (r) *t = tt;
    if (margin) *margin = ctx.margin;
    return r;
}
int orc_edge_edge_ctcd(const double *q, double eta, double *t, double *margin)
{
    Ctx ctx;
    double tt = 0;
    bool r = edge_edge_ctcd(load3(q), load3(q + 3), load3(q + 6), load3(q + 9), load3(q + 12), load3(q + 15),
                            load3(q + 18), load3(q + 21), eta, tt, &ctx);
    if (r) *t = tt;
    if (margin) *margin = ctx.margin;
    return r;
}
int orc_vertex_edge_ctcd(const double *q, double eta, double *t, double *margin)
{
    Ctx ctx;
    double tt = 0;
    bool r = vertex_edge_ctcd(load3(q), load3(q + 3), load3(q + 6), load3(q + 9), load3(q + 12), load3(q + 15),
                              eta, tt, &ctx);
    if (r) *t = tt;
    if (margin) *margin = ctx.margin;
    return r;
}
int orc_vertex_vertex_ctcd(const double *q, double eta, double *t, double *margin)
{
    Ctx ctx;
    double tt = 0;
    bool r = vertex_vertex_ctcd(load3(q), load3(q + 3), load3(q + 6), load3(q + 9), eta, tt, &ctx);
    if (r) *t = tt;
    if (margin) *margin = ctx.margin;
    return r;
}

// Leaf bounds (13 mins then 13 maxs per face; AABB uses the first 3 of each).
void orc_leaf_dops(const int32_t *faces, int64_t F, const int64_t *off, const double *ht, const double *hp,
                   double eta, int aabb, double *out_mn, double *out_mx)
{
    Hist h{off, ht, hp};
    std::vector<Dop> d;
    leaf_dops(faces, F, h, eta, aabb != 0, d);
    for (int64_t f = 0; f < F; f++)
        for (int k = 0; k < KAX; k++) {
            out_mn[KAX * f + k] = d[f].mn[k];
            out_mx[KAX * f + k] = d[f].mx[k];
        }
}

// Candidate stencil sets.  method: 0 = brute force, 1 = reference-structured tree with std::set,
// 2 = tree with vector + sort/unique.  aabb selects AABBBroadPhase's 3-axis bounds.
void *orc_candidates(const int32_t *faces, int64_t F, const int64_t *off, const double *ht, const double *hp,
                     double eta, const uint8_t *fixed, int method, int aabb, int want_pairs)
{
    Hist h{off, ht, hp};
    std::vector<Dop> d;
    leaf_dops(faces, F, h, eta, aabb != 0, d);
    Result *r = new Result;
    int na = aabb ? 3 : KAX;
    if (method == 0)
        candidates_bruteforce(faces, F, d, na, fixed, r->cs, want_pairs ? &r->pairs : nullptr);
    else
        candidates_tree(faces, F, d, na, fixed, r->cs, method == 1, method == 3);
    return r;
}
int64_t orc_result_count(void *r, int which)
{
    Result *R = (Result *)r;
    if (which == 0) return (int64_t)R->cs.vf.size();
    if (which == 1) return (int64_t)R->cs.ee.size();
    return (int64_t)R->pairs.size() / 2;
}
void orc_result_copy(void *r, int which, int32_t *out)
{
    Result *R = (Result *)r;
    if (which == 2) {
        for (size_t i = 0; i < R->pairs.size(); i++) out[i] = (int32_t)R->pairs[i];
        return;
    }
    const std::vector<Key> &v = which == 0 ? R->cs.vf : R->cs.ee;
    std::memcpy(out, v.data(), v.size() * sizeof(Key));
}
void orc_result_free(void *r) { delete (Result *)r; }

// Narrowphase over stencil lists (CTCDNarrowPhase.cpp:9-22), one record per stencil.  is_ee selects
// checkEES.  counts (optional) receives the ST_COUNT work counters summed over all stencils.
void orc_narrowphase(const int64_t *off, const double *ht, const double *hp, const int32_t *keys, const double *eta,
                     int64_t n, int is_ee, void *out_records, int64_t *counts)
{
    Hist h{off, ht, hp};
    StencilResult *R = (StencilResult *)out_records;
    int64_t total[ST_COUNT] = {};
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) reduction(+ : total[:ST_COUNT])
    for (int64_t i = 0; i < n; i++) {
        Ctx ctx;
        StencilResult res;
        res.hit = 0;
        res.subtest = -1;
        res.window = -1;
        res.pad = 0;
        res.t = 0;
        bool hit = is_ee ? check_ees(h, keys + 4 * i, eta[i], res, &ctx) : check_vfs(h, keys + 4 * i, eta[i], res, &ctx);
        res.hit = hit ? 1 : 0;
        if (!hit) {
            res.subtest = -1;
            res.window = -1;
            res.t = 0;
        }
        res.margin = ctx.margin;
        R[i] = res;
        for (int k = 0; k < ST_COUNT; k++) total[k] += ctx.counts[k];
    }
    if (counts)
        for (int k = 0; k < ST_COUNT; k++) counts[k] += total[k];
}

// SeparatingPlaneNarrowPhase::findCollisions (SeparatingPlaneNarrowPhase.cpp:9-24) per stencil: hit
// 1 / 0, or 2 where the reference's recursion would not terminate; the decision margin of the
// primitive tests run; the deepest recursion level reached.
double orc_minimum_gap(const int64_t *off, const double *ht, const double *hp, int64_t nverts)
{
    return compute_minimum_gap(Hist{off, ht, hp}, nverts);
}

void orc_separating_plane(const int64_t *off, const double *ht, const double *hp, int64_t nverts, const int32_t *keys,
                          const double *eta, int64_t n, int is_ee, int32_t *hit, double *margin, int32_t *depth)
{
    Hist h{off, ht, hp};
    double eps = compute_minimum_gap(h, nverts) / 4.0;
    for (int64_t i = 0; i < n; i++) {
        Ctx ctx;
        int verts[4] = {keys[4 * i], keys[4 * i + 1], keys[4 * i + 2], keys[4 * i + 3]};
        int md = 0, und = 0;
        bool r = sp_check_interval(!is_ee, h, verts, eta[i], 0.0, 1.0, eps, &ctx, 0, md, &und);
        hit[i] = und ? 2 : (r ? 1 : 0);
        if (margin) margin[i] = ctx.margin;
        if (depth) depth[i] = md;
    }
}

int orc_stage_count() { return ST_COUNT; }

// Proximity statics on one item: q = 4 points (12 doubles).  kind as CCD_DIST_* in include/ccd_amd.h.
// Smallest distance between each stencil's primitives (vertex-face or edge-edge, Distance.h:32-174)
// over `samples` equally spaced times of the straight motion qs -> qe: a check of the narrowphases
// (which must report every stencil that comes within eta), independent of their polynomials.
void orc_sampled_min_distance(const double *qs, const double *qe, const int32_t *keys, int64_t n, int is_ee,
                              int samples, double *out)
{
    for (int64_t i = 0; i < n; i++) {
        double best = std::numeric_limits<double>::infinity();
        for (int k = 0; k < samples; k++) {
            const double t = samples > 1 ? (double)k / (double)(samples - 1) : 0.0;
            V3 P[4];
            for (int j = 0; j < 4; j++) {
                const V3 a = load3(qs + 3 * (int64_t)keys[4 * i + j]), b = load3(qe + 3 * (int64_t)keys[4 * i + j]);
                P[j] = V3{(1 - t) * a.x + t * b.x, (1 - t) * a.y + t * b.y, (1 - t) * a.z + t * b.z};
            }
            double bary[4];
            const V3 d = is_ee ? edge_edge_distance(P[0], P[1], P[2], P[3], bary)
                               : vertex_face_distance(P[0], P[1], P[2], P[3], bary);
            best = std::min(best, norm(d));
        }
        out[i] = best;
    }
}

int orc_distance(int kind, const double *q, double eta, double *vec, double *bary)
{
    V3 a = load3(q), b = load3(q + 3), c = load3(q + 6), d = load3(q + 9);
    if (kind == 0) return vertex_plane_lt(a, b, c, d, eta);
    if (kind == 1) return line_line_lt(a, b, c, d, eta);
    V3 r = kind == 2 ? vertex_face_distance(a, b, c, d, bary) : edge_edge_distance(a, b, c, d, bary);
    vec[0] = r.x;
    vec[1] = r.y;
    vec[2] = r.z;
    return 1;
}

double orc_mesh_self_distance(const double *verts, int64_t V, const int32_t *faces, int64_t F, const uint8_t *fixed,
                              int64_t *nvf, int64_t *nee)
{
    return mesh_self_distance(verts, V, faces, F, fixed, nvf, nee);
}

// VelocityFilter::velocityFilter on host arrays; qend is overwritten on success.  stats receives
// iterations, layers, firings, history entries of the last iteration, VF / EE candidates and collisions
// summed over iterations, undecided stencils (9 int64).  *hist (optional) receives the final iteration's
// history on success (free with orc_vhist_free).
int orc_velocity_filter(const double *qstart, double *qend, int64_t V, const int32_t *faces, int64_t F,
                        const double *invm, double outer, double inner, double stiff, double substep, double cor,
                        int max_rollbacks, int64_t *stats, double *closest, void **hist)
{
    VFStats st;
    VHist *h = nullptr;
    int r = velocity_filter(qstart, qend, V, faces, F, invm, outer, inner, stiff, substep, cor, max_rollbacks, st,
                            hist ? &h : nullptr);
    if (stats) {
        int64_t v[9] = {st.iterations, st.layers, st.firings, st.history_entries, st.vf_candidates,
                        st.ee_candidates, st.vf_collisions, st.ee_collisions, st.undecided};
        std::memcpy(stats, v, sizeof v);
    }
    if (closest) *closest = st.closest;
    if (hist) *hist = h;
    return r;
}
int64_t orc_vhist_entries(void *h) { return ((VHist *)h)->entries(); }
void orc_vhist_copy(void *h, int64_t *off, double *t, double *p)
{
    std::vector<int64_t> o;
    std::vector<double> tt, pp;
    ((VHist *)h)->csr(o, tt, pp);
    std::memcpy(off, o.data(), o.size() * sizeof(int64_t));
    std::memcpy(t, tt.data(), tt.size() * sizeof(double));
    std::memcpy(p, pp.data(), pp.size() * sizeof(double));
}
void orc_vhist_free(void *h) { delete (VHist *)h; }

}   // extern "C"
