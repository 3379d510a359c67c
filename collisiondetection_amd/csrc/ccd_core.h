// FP64 core of the swept-primitive CCD engine: the conservative eta-CCD primitive tests of
// evouga/collisiondetection (src/CTCD.cpp) and the stencil-level narrowphase of src/CTCDNarrowPhase.cpp,
// written for one GPU lane per stencil.
//
// Numerics.  Everything the reference decides with is reproduced operation by operation, so this
// file must be compiled with -ffp-contract=off (no fused multiply-add may be formed from a*b+c):
//   * dot products are summed ((x0*y0 + x1*y1) + x2*y2), cross products are
//     (a1*b2 - a2*b1, a2*b0 - a0*b2, a0*b1 - a1*b0) -- the arithmetic of Eigen's Vector3d;
//   * every polynomial coefficient follows the reference's expression order (CTCD.cpp:179-257,
//     :536-570, :616-621), is normalised by division by its largest magnitude and degree-reduced
//     exactly as CTCD.cpp:113-133 does;
//   * degree <= 2 polynomials use the reference's closed forms (CTCD.cpp:38-56, :143-158), so their
//     intervals are bit-identical, NaN corner cases included;
//   * the sign of each piece between roots is decided at its midpoint with plain Horner on the
//     normalised coefficients (CTCD.cpp:59-79), and std::max / std::min keep their (a < b) ? b : a
//     and (b < a) ? b : a argument order.
// Roots of degree 3..6 polynomials are isolated, not found by the reference's Jenkins-Traub solver
// (rpoly.h): the real roots in [0,1] come from recursion on derivatives (critical points split [0,1]
// into monotone pieces; each piece with a sign change holds exactly one root, found by bracketed
// Newton iteration).  The real parts of complex roots, which the reference also uses as split points,
// never change the union of the sign intervals, so vertex-face and vertex-edge results agree except
// where a root lies within rounding distance of a decision.  Times of impact are the exact roots of
// the double polynomials to ~1e-15, where the reference's solver can be off by far more (its error on
// the near-double roots of the distance sextics reaches 1e-4).
#pragma once

#include <stdint.h>

#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
// The device cascade is inlined into its kernels: its interval lists and polynomials then stay in
// registers instead of the stack frames of out-of-line calls (per-lane scratch).  A file that defines
// CCD_OUTLINE_CASCADE leaves the inlining to the compiler.
#if defined(CCD_OUTLINE_CASCADE)
#define CCD_HD __host__ __device__ inline
#else
#define CCD_HD __host__ __device__ __forceinline__
#endif
#define CCD_FHD __host__ __device__ __forceinline__
#define CCD_HD_NI __host__ __device__   // with static __attribute__((noinline)): one out-of-line copy per file
#else
#include <cmath>
#define CCD_HD inline
#define CCD_FHD inline
#define CCD_HD_NI inline
#endif

// Test builds for the host (CCD_CHECK_INDEX) stop on any runtime index outside its array; the product
// build compiles the checks away.
#if defined(CCD_CHECK_INDEX) && !defined(__HIPCC__)
#include <cstdio>
#include <cstdlib>
#define CCD_INDEX_CHECK(i, n)                                                                                      \
    do {                                                                                                           \
        if ((i) < 0 || (i) >= (n)) {                                                                               \
            std::fprintf(stderr, "ccd_core: index %d outside [0, %d) at line %d\n", (int)(i), (int)(n), __LINE__); \
            std::abort();                                                                                          \
        }                                                                                                          \
    } while (0)
#else
#define CCD_INDEX_CHECK(i, n) ((void)0)
#endif

namespace ccd {

// ------------------------------------------------------------------------------------------------
// Scalars and 3-vectors
// ------------------------------------------------------------------------------------------------
CCD_FHD double smax(double a, double b) { return (a < b) ? b : a; }   // std::max(a, b)
CCD_FHD double smin(double a, double b) { return (b < a) ? b : a; }   // std::min(a, b)
CCD_FHD double dabs(double a) { return __builtin_fabs(a); }
CCD_FHD double dsqrt(double a)
{
#if defined(__HIPCC__)
    return __builtin_sqrt(a);
#else
    return std::sqrt(a);
#endif
}
CCD_FHD double dfma(double a, double b, double c)
{
#if defined(__HIPCC__)
    return __builtin_fma(a, b, c);
#else
    return std::fma(a, b, c);
#endif
}

struct V3 {
    double x, y, z;
};
CCD_FHD V3 mk(double x, double y, double z)
{
    V3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
CCD_FHD V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
CCD_FHD V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
CCD_FHD V3 operator*(double s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
CCD_FHD double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
CCD_FHD V3 cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
CCD_FHD double norm(V3 a) { return dsqrt(dot(a, a)); }

// ------------------------------------------------------------------------------------------------
// Work accounting.  A counter functor is called once per executed stage; NoCount compiles away.
// The FP64 operations (add, sub, mul, div, sqrt) each stage performs to build its coefficients are
// listed in STAGE_FLOPS; root isolation is not counted (it depends on the method).
// ------------------------------------------------------------------------------------------------
enum Stage {
    S_VF_CALL = 0,   // vertexFaceCTCD entry: 4 velocities                                   (CTCD.cpp:423-427)
    S_VF_PLANE,      // one edge-side cubic: its 6 vectors, 4 coefficients, normalisation   (:433-464, :215-229)
    S_VF_DIST,       // the coplanar-distance sextic                                        (:466-475, :231-257)
    S_EE_DIST,       // edgeEdgeCTCD entry: velocities and the distance sextic              (:270-288)
    S_EE_PAR,        // one parallel-edge check of a raw coplanar interval                  (:294-308)
    S_EE_BARY,       // one barycentric quartic                                             (:313-350, :179-213)
    S_VE_CALL,       // vertexEdgeCTCD entry: velocities and the 6 edge vectors             (:520-534)
    S_VE_QUAD,       // one side quadratic                                                  (:536-555)
    S_VE_QUART,      // the colinear-distance quartic                                       (:557-573)
    S_VV,            // vertexVertexCTCD's quadratic coefficients                            (:610-621)
    S_QUAD_ROOT,     // one getQuadRoots call with real roots                                (:38-56)
    S_ISOLATE,       // one polynomial of reduced degree > 2 sent to root isolation          (:135-142)
    S_COUNT
};
// Flops per stage (add/sub/mul/div/sqrt = 1 each), counted from the expressions below.
constexpr int STAGE_FLOPS[S_COUNT] = {12, 126, 172, 184, 41, 130, 27, 19, 93, 68, 10, 0};

struct NoCount {
    CCD_FHD void operator()(int) const {}
};

// The product's counter types prune cascades that cannot hit (see any_overlap below); the counting
// types used for work accounting, and FullCascade, run the reference's full cascade.
struct FullCascade {
    CCD_FHD void operator()(int) const {}
};
// The product's counter type for the first pass of the two-pass edge-edge narrowphase: it screens the
// coplanar-distance sextic and solves nothing.  find_intervals on the sextic decides only the case
// "full degree, and no root in [0, 1] by the Bernstein test" (couldHaveRoots false, or no sign change
// of the Bernstein coefficients with every one clear of zero), where the list is the one piece [0, 1]
// or nothing, as the isolator gives it; any other sextic sets `deferred`.  edge_edge then returns
// false at once, and also sets `deferred` when the sextic leaves a coplanar interval.  So a stencil
// the screen does not defer has no edge-edge collision, exactly as the full test finds; the deferred
// ones get the full test later, on a list that holds only them.  The first pass then carries no root
// solver, no parallel check and no barycentric quartics for the edge-edge test (fewer registers: more
// wavefronts per SIMD), and the stencils whose sextics need the solver run together.
struct ScreenSextic {
    bool deferred = false;
    CCD_FHD void operator()(int) const {}
};
template <class C>
struct prunes : std::integral_constant<bool, std::is_same<std::decay_t<C>, NoCount>::value ||
                                                 std::is_same<std::decay_t<C>, ScreenSextic>::value> {};
template <class C>
struct screens_sextic : std::integral_constant<bool, std::is_same<std::decay_t<C>, ScreenSextic>::value> {};
// A counter type whose `stops_at_solver` member is true makes find_intervals return, with no intervals,
// where it would call the root solver.  SolverProbe records whether that point was reached; the
// stencil-ordering pre-pass uses it to see which polynomials need the solver.
template <class C, class = void>
struct stops_at_solver : std::false_type {};
template <class C>
struct stops_at_solver<C, std::void_t<decltype(C::stops_at_solver)>> : std::integral_constant<bool, C::stops_at_solver> {};

// A counter type whose `plain_roots` member is true isolates roots without the Bernstein shortcuts and
// with plain safeguarded Newton (bisection fallback, plain Horner values): the work-accounting kernel.
template <class C, class = void>
struct plain_roots : std::false_type {};
template <class C>
struct plain_roots<C, std::void_t<decltype(C::plain_roots)>> : std::integral_constant<bool, C::plain_roots> {};

CCD_FHD double bits_to_double(long long v)
{
    double d;
    __builtin_memcpy(&d, &v, 8);
    return d;
}
CCD_FHD long long double_to_bits(double d)
{
    long long v;
    __builtin_memcpy(&v, &d, 8);
    return v;
}

// Static-index access.  On the GPU an array indexed by a runtime value lives in scratch memory (per-lane
// private memory, cached but far slower than registers, and written back to memory).  Every list here
// is small, so reads and writes at a runtime index go through selects and the arrays stay in registers.
// The selects are spelled out by template recursion, so every element access has a constant index
// from the start (a loop over the elements still has a variable index when arrays are first promoted
// to registers), and a write assigns every element (a conditional store per element would be merged
// back into one store at a runtime address).
template <int K, int N>
CCD_FHD double sel_from(const double (&a)[N], int i)
{
    if constexpr (K == N - 1)
        return a[K];
    else
        return i == K ? a[K] : sel_from<K + 1>(a, i);
}
template <int N>
CCD_FHD double sel(const double (&a)[N], int i)
{
    CCD_INDEX_CHECK(i, N);
    return sel_from<0>(a, i);
}
template <int K, int N>
CCD_FHD void put_from(double (&a)[N], int i, double v)
{
    a[K] = i == K ? v : a[K];
    if constexpr (K + 1 < N) put_from<K + 1>(a, i, v);
}
template <int N>
CCD_FHD void put(double (&a)[N], int i, double v)
{
    CCD_INDEX_CHECK(i, N);
    put_from<0>(a, i, v);
}

// checkInterval's clamp of an interval end to [0, 1] (CTCD.cpp:62-65), in its operation order.
CCD_FHD double clamp_cut(double t) { return smin(1.0, smax(0.0, t)); }

// ------------------------------------------------------------------------------------------------
// Interval lists (CTCD.h:7-26).  Every list findIntervals builds is a subset of the pieces of one
// partition of [0, 1], in ascending order: the partition's cut points are the polynomial's roots, each
// clamped to [0, 1] as checkInterval clamps it, and the list holds the pieces that pass the midpoint
// sign test.  So a list is kept as its nb cut points c[0..nb) and a bit mask of the pieces it holds:
// piece p runs from cut p-1 (0 for p = 0) to cut p (1 for p = nb), and its TimeInterval is those two
// ends ordered and clamped as the constructor does (CTCD.h:9-14).  The i-th interval of the reference's
// list is the i-th piece in `keep`.  CAP pieces need CAP - 1 cut points, so a list takes about half the
// registers of (lower, upper) pairs.
// ------------------------------------------------------------------------------------------------
template <int CAP>
// Every slot starts defined: the cut points are read through unrolled selects whose guards the
// compiler may evaluate together with the read.
struct Ivls {
    double c[CAP > 1 ? CAP - 1 : 1] = {};
    int nb = 0;
    unsigned keep = 0;

    CCD_FHD bool empty() const { return keep == 0; }
    // The piece's ends as checkInterval sees them (already clamped, in the order given).
    CCD_FHD double cut_lo(int p) const { return p == 0 ? 0.0 : sel(c, p - 1); }
    CCD_FHD double cut_hi(int p) const { return p == nb ? 1.0 : sel(c, p); }
    // TimeInterval(cut_lo, cut_hi): ordered, then clamped (CTCD.h:9-14).
    CCD_FHD void get(int p, double &l, double &u) const
    {
        l = cut_lo(p);
        u = cut_hi(p);
        if (l > u) {
            double s = l;
            l = u;
            u = s;
        }
        l = smax(l, 0.0);
        u = smin(u, 1.0);
    }
    CCD_FHD double lo(int p) const
    {
        double l, u;
        get(p, l, u);
        return l;
    }
    CCD_FHD double hi(int p) const
    {
        double l, u;
        get(p, l, u);
        return u;
    }
};

// Lowest set bit of a nonzero mask.
CCD_FHD int low_bit(unsigned m) { return __builtin_ctz(m); }

// Horner in the reference's order (CTCD.cpp:68-73), no fused multiply-add.
template <int D>
CCD_FHD double horner(const double *op, int degree, double x)
{
    double f = op[0];
#pragma unroll
    for (int i = 1; i <= D; i++) {
        if (i <= degree) {
            f *= x;
            f += op[i];
        }
    }
    return f;
}

// checkInterval's sign test (CTCD.cpp:59-79) on ends that are already clamped.
template <int D>
CCD_FHD bool midpoint_passes(double t1, double t2, const double *op, int degree, bool pos)
{
    double tmid = (t2 + t1) / 2;
    double f = horner<D>(op, degree, tmid);
    return (pos && f >= 0) || (!pos && f <= 0);
}

// checkInterval (CTCD.cpp:59-79) for piece p of `out`: keeps it if its midpoint passes.
template <int D, int CAP>
CCD_FHD void check_piece(int p, const double *op, int degree, Ivls<CAP> &out, bool pos)
{
    if (midpoint_passes<D>(out.cut_lo(p), out.cut_hi(p), op, degree, pos)) out.keep |= 1u << p;
}

// checkInterval for an interval that is not kept, only tested (vertexVertexCTCD's pieces).
template <int D>
CCD_FHD bool check_interval_passes(double t1, double t2, const double *op, int degree, bool pos)
{
    t1 = smax(0.0, t1);
    t2 = smax(0.0, t2);
    t1 = smin(1.0, t1);
    t2 = smin(1.0, t2);
    return midpoint_passes<D>(t1, t2, op, degree, pos);
}

// couldHaveRoots (CTCD.cpp:81-94) on op[0..degree], degree <= D.
template <int D>
CCD_FHD bool could_have_roots(const double *op, int degree, bool pos)
{
    double result = 0;
    if ((pos && op[0] > 0) || (!pos && op[0] < 0)) result = op[0];
    double last = op[0];
#pragma unroll
    for (int i = 1; i <= D; i++) {
        if (i < degree) {
            result *= 1.0;
            if ((pos && op[i] > 0) || (!pos && op[i] < 0)) result += op[i];
        }
        if (i == degree) last = op[i];
    }
    result *= 1.0;
    result += last;
    return !((pos && result < 0) || (!pos && result > 0));
}

// getQuadRoots (CTCD.cpp:38-56)
CCD_FHD int quad_roots(double a, double b, double c, double &t0, double &t1)
{
    int roots = 0;
    int sign = 1;
    if (b < 0) sign = -1;
    double D = b * b - 4 * a * c;
    if (D >= 0) {
        roots = 2;
        double q = -0.5 * (b + sign * dsqrt(D));
        t0 = q / a;
        t1 = c / q;
        if (t0 > t1) {
            double s = t0;
            t0 = t1;
            t1 = s;
        }
    }
    return roots;
}

// ------------------------------------------------------------------------------------------------
// Real-root isolation on [0,1] for degree 1..6 (used for reduced degree >= 3 only).
// c[0..D] descending, c[0] != 0.  Writes the distinct real roots that lie in [0,1], ascending, and
// returns their count.  A root exactly at an end (c evaluates to exactly 0 there) is reported as
// exactly that end.  Polynomials whose Bernstein coefficients show no or one sign change are settled
// directly; the others recurse on the derivative.
// ------------------------------------------------------------------------------------------------
template <int D>
CCD_FHD double peval(const double *c, double x)
{
    double f = c[0];
#pragma unroll
    for (int i = 1; i <= D; i++) f = dfma(f, x, c[i]);
    return f;
}
template <int D>
CCD_FHD void peval_d(const double *c, double x, double &f, double &df)
{
    f = c[0];
    df = 0.0;
#pragma unroll
    for (int i = 1; i <= D; i++) {
        df = dfma(df, x, f);
        f = dfma(f, x, c[i]);
    }
}

// Value, first and second derivative of c at x (Horner with fused multiply-add).
template <int D>
CCD_FHD void peval_d2(const double *c, double x, double &f, double &df, double &d2)
{
    f = c[0];
    df = 0.0;
    d2 = 0.0;
#pragma unroll
    for (int i = 1; i <= D; i++) {
        d2 = dfma(d2, x, df);
        df = dfma(df, x, f);
        f = dfma(f, x, c[i]);
    }
    d2 *= 2.0;
}

// Value and derivative of c at x, and m = sum |c_i| |x|^(D-i): the value's rounding error is a small
// multiple of u m, and below u m its sign is no longer reliable.
template <int D>
CCD_FHD void peval_dm(const double *c, double x, double &f, double &df, double &m)
{
    f = c[0];
    df = 0.0;
    m = dabs(c[0]);
    const double ax = dabs(x);
#pragma unroll
    for (int i = 1; i <= D; i++) {
        df = dfma(df, x, f);
        f = dfma(f, x, c[i]);
        m = dfma(m, ax, dabs(c[i]));
    }
}

// c at x by compensated Horner: each product's and sum's rounding error is recovered exactly (the
// product's with a fused multiply-add, the sum's with Knuth's two-sum) and accumulated in a second
// Horner pass, so the value is about as accurate as Horner in twice the precision (Graillat, Langlois
// and Louvet, "Compensated Horner scheme", 2005).
template <int D>
CCD_FHD double peval_comp(const double *c, double x)
{
    double s = c[0], e = 0.0;
#pragma unroll
    for (int i = 1; i <= D; i++) {
        const double p = s * x;
        const double pe = dfma(s, x, -p);
        const double sn = p + c[i];
        const double t = sn - p;
        const double se = (p - (sn - t)) + (c[i] - t);
        e = dfma(e, x, pe + se);
        s = sn;
    }
    return s + e;
}

// The single root of c in [a, b] given opposite nonzero signs fa, fb at the ends: safeguarded Newton
// (a step that leaves the bracket becomes a bisection).
//   * Start: the bracket ends are critical points of c (or 0 and 1), and near a pair of close roots --
//     which the distance polynomials have, about eta apart -- c is locally quadratic there: Newton from
//     the far side would close in only linearly.  So the start is the root of the second-order Taylor
//     model at the end with the smaller |c| (the one nearest it, towards the other end), else the
//     secant point.
//   * Once the plain value falls below its rounding level (u sum |c_i| |x|^(D-i)), the value comes from
//     compensated Horner, so the iteration keeps converging instead of following rounding noise; the
//     derivative stays plain (it only sets the step).
//   * Stop when the Newton step is below the resolution of x, the bracket has closed, or c is exactly 0.
template <int D, bool PLAIN = false>
CCD_HD double bracketed_root(const double *c, double a, double b, double fa, double fb)
{
    if constexpr (PLAIN) {
        // Newton from the secant point, bisection whenever a step leaves the bracket.
        double lo = a, hi = b;
        const bool neg_lo = fa < 0;
        double x = a + (b - a) * (fa / (fa - fb));
        if (!(x > lo && x < hi)) x = 0.5 * (lo + hi);
        for (int it = 0; it < 80; it++) {
            double f, df;
            peval_d<D>(c, x, f, df);
            if (f == 0.0) return x;
            if ((f < 0) == neg_lo)
                lo = x;
            else
                hi = x;
            double xn = x - f / df;
            if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
            double step = xn - x;
            if (step < 0) step = -step;
            x = xn;
            const double ax = x < 0 ? -x : x;
            if (step <= 2.3e-16 * ax + 1e-18 || hi - lo <= 2.3e-16 * ax + 1e-18 || !(hi > lo)) break;
        }
        return x;
    }
    double lo = a, hi = b;
    const bool neg_lo = fa < 0;
    double x = a + (b - a) * (fa / (fa - fb));
    {
        const bool at_a = dabs(fa) <= dabs(fb);
        const double e = at_a ? a : b, dir = at_a ? 1.0 : -1.0;
        double p, p1, p2;
        peval_d2<D>(c, e, p, p1, p2);
        // e + dir u with u > 0 the smallest root of p + (p1 dir) u + (p2 / 2) u^2
        const double A = 0.5 * p2, B = p1 * dir, C = p;
        double u = -1.0;
        if (A != 0.0) {
            const double disc = B * B - 4.0 * A * C;
            if (disc >= 0.0) {
                const double sq = dsqrt(disc);
                const double q = -0.5 * (B + (B < 0 ? -sq : sq));
                const double u1 = q / A, u2 = q != 0.0 ? C / q : -1.0;
                if (u1 > 0.0) u = u1;
                if (u2 > 0.0 && (u < 0.0 || u2 < u)) u = u2;
            }
        } else if (B != 0.0) {
            u = -C / B;
        }
        if (u > 0.0) {
            const double x0 = e + dir * u;
            if (x0 > lo && x0 < hi) x = x0;
        }
    }
    if (!(x > lo && x < hi)) x = 0.5 * (lo + hi);
    bool compensated = false;
    for (int it = 0; it < 100; it++) {
        double f, df, m;
        peval_dm<D>(c, x, f, df, m);
        if (!compensated && dabs(f) <= 1.1102230246251565e-16 * m) compensated = true;
        if (compensated) f = peval_comp<D>(c, x);
        if (f == 0.0) return x;
        if ((f < 0) == neg_lo)
            lo = x;
        else
            hi = x;
        const double step = f / df;
        const double ax = x < 0 ? -x : x;
        if (dabs(step) <= 2.3e-16 * ax + 1e-18) return x - step;
        double xn = x - step;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        x = xn;
        const double axn = x < 0 ? -x : x;
        if (hi - lo <= 2.3e-16 * axn + 1e-18 || !(hi > lo)) break;
    }
    return x;
}

// Binomial coefficient as a double (exact for the small arguments used here); constant-folded.
CCD_FHD constexpr double binom(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (double)(n - k + i) / (double)i;
    return r;
}

// Sign variations of the Bernstein coefficients of c[0..D] (descending powers) on [0, 1], or -1 when one
// of them lies within a margin of zero that covers the rounding of the conversion.  p(t) on [0, 1] is a
// convex combination of its Bernstein coefficients b_k = sum_{i <= k} C(k, i) / C(D, i) a_i (a_i the
// coefficient of t^i), and the number of roots in (0, 1) is at most the number of sign changes of b and
// has its parity (Descartes' rule of signs in the Bernstein basis).  So with every b_k clear of zero, no
// variation means no root in [0, 1] and one variation exactly one root, inside (0, 1).  The margin:
// each computed b_k is off by at most (k + 2) u sum |a_i| <= 56 u max |a_i| (D <= 6, u = 2^-53), and
// 1e-13 max |a_i| is 16 times that.  Non-finite coefficients give -1.
template <int D>
CCD_FHD int bernstein_variations(const double *c)
{
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i <= D; i++) mx = smax(mx, dabs(c[i]));
    const double tol = 1e-13 * mx;
    int v = 0, last = 0;
    bool clear = true;
#pragma unroll
    for (int k = 0; k <= D; k++) {
        double b = 0.0;
#pragma unroll
        for (int i = 0; i <= k; i++) b += (binom(k, i) / binom(D, i)) * c[D - i];
        const int sg = b > tol ? 1 : (b < -tol ? -1 : 0);
        if (sg == 0) clear = false;
        if (sg != 0 && last != 0 && sg != last) v++;
        if (sg != 0) last = sg;
    }
    return clear ? v : -1;
}

template <int D, bool PLAIN = false>
struct RootIsolator {
    CCD_HD static int roots01(const double *c, double (&r)[D])
    {
        // The common cases first (most polynomials the cascade solves): no root in [0, 1], or exactly one
        // in (0, 1), bracketed by the ends, where p(0) = c[D] exactly and p(1) has b_D's sign.
        if constexpr (!PLAIN) {
            const int v = bernstein_variations<D>(c);
            if (v == 0) return 0;
            if (v == 1) {
                put(r, 0, bracketed_root<D>(c, 0.0, 1.0, peval<D>(c, 0.0), peval<D>(c, 1.0)));
                return 1;
            }
        }
        double d[D];
#pragma unroll
        for (int i = 0; i < D; i++) d[i] = c[i] * (double)(D - i);
        double cr[D - 1] = {};
        int nc = RootIsolator<D - 1, PLAIN>::roots01(d, cr);
        int nr = 0;
        double a = 0.0;
        double fa = peval<D>(c, 0.0);
        double last = 0.0;   // r[nr - 1]
        if (fa == 0.0) {
            put(r, nr++, 0.0);
            last = 0.0;
        }
#pragma unroll
        for (int k = 0; k < D; k++) {
            if (k <= nc) {
                double b = (k < nc) ? cr[k] : 1.0;
                if (b > a) {
                    double fb = peval<D>(c, b);
                    if (fb == 0.0) {
                        if (nr == 0 || last != b) {
                            put(r, nr++, b);
                            last = b;
                        }
                    } else if (fa != 0.0 && ((fa < 0) != (fb < 0))) {
                        last = bracketed_root<D, PLAIN>(c, a, b, fa, fb);
                        put(r, nr++, last);
                    }
                    a = b;
                    fa = fb;
                }
            }
        }
        return nr;
    }
};
template <bool PLAIN>
struct RootIsolator<1, PLAIN> {
    CCD_HD static int roots01(const double *c, double (&r)[1])
    {
        double x = -c[1] / c[0];
        if (x >= 0.0 && x <= 1.0) {
            r[0] = x == 0.0 ? 0.0 : x;
            return 1;
        }
        return 0;
    }
};
template <bool PLAIN>
struct RootIsolator<2, PLAIN> {
    CCD_HD static int roots01(const double *c, double (&r)[2])
    {
        double a = c[0], b = c[1], cc = c[2];
        double disc = b * b - 4.0 * a * cc;
        if (disc < 0) return 0;
        double sq = dsqrt(disc);
        double q = -0.5 * (b + (b < 0 ? -sq : sq));
        double x0, x1;
        if (q == 0.0) {
            x0 = x1 = -b / (2.0 * a);
        } else {
            x0 = q / a;
            x1 = cc / q;
        }
        if (x0 > x1) {
            double s = x0;
            x0 = x1;
            x1 = s;
        }
        int n = 0;
        if (x0 >= 0.0 && x0 <= 1.0) {
            r[0] = x0;
            n = 1;
        }
        if (x1 >= 0.0 && x1 <= 1.0 && !(n > 0 && x0 == x1)) {
            if (n == 0)
                r[0] = x1;
            else
                r[1] = x1;
            n++;
        }
        return n;
    }
};

// Roots of the degree-D polynomial op in [0,1] into time[0..), D <= M.
template <int D, int M, bool PLAIN = false>
CCD_HD int isolate_into(const double *op, double (&time)[M])
{
    double r[D] = {};
    int n = RootIsolator<D, PLAIN>::roots01(op, r);
#pragma unroll
    for (int k = 0; k < D && k < M; k++) time[k] = r[k];
    return n;
}

// Real roots in [0,1] of op[0..degree] (3 <= degree <= N) by isolation.
template <int N, int M, bool PLAIN = false>
CCD_FHD int isolate_roots(const double *op, int degree, double (&time)[M])
{
    int roots = 0;
    if (degree == 3) roots = isolate_into<3, M, PLAIN>(op, time);
    if (N >= 4 && degree == 4) roots = isolate_into<(N >= 4 ? 4 : 3), M, PLAIN>(op, time);
    if (N >= 5 && degree == 5) roots = isolate_into<(N >= 5 ? 5 : 3), M, PLAIN>(op, time);
    if (N >= 6 && degree == 6) roots = isolate_into<(N >= 6 ? 6 : 3), M, PLAIN>(op, time);
    return roots;
}

// ------------------------------------------------------------------------------------------------
// findIntervals (CTCD.cpp:98-177) for a degree-N polynomial, op modified in place.
// ------------------------------------------------------------------------------------------------
template <int N, int CAP, class C = NoCount>
CCD_HD void find_intervals(double *op, Ivls<CAP> &out, bool pos, C &&cnt = C())
{
    static_assert(CAP >= (N > 2 ? N : 2) + 1, "an interval list holds every piece of its partition");
    using CT = std::decay_t<C>;
    out.nb = 0;
    out.keep = 0;
    double maxval = 0;
#pragma unroll
    for (int i = 0; i <= N; i++) maxval = smax(maxval, dabs(op[i]));
    if (maxval != 0) {
#pragma unroll
        for (int i = 0; i <= N; i++) op[i] /= maxval;
    }
    int reduced = N;
#pragma unroll
    for (int i = 0; i < N; i++) {
        if (reduced == N - i && op[i] == 0) reduced--;
    }
    if (reduced < N) {
        int s = N - reduced;
        // op[i] = op[i + s] for i = 0..reduced, with unrolled static indexing.
        double tmp[N + 1];
#pragma unroll
        for (int i = 0; i <= N; i++) tmp[i] = op[i];
#pragma unroll
        for (int i = 0; i <= N; i++) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k <= N; k++)
                if (k == i + s) v = tmp[k];
            if (i <= reduced) op[i] = v;
        }
    }

    if constexpr (screens_sextic<CT>::value && N == 6) {
        if (reduced <= 2) {   // a degree-reduced sextic: not screened
            cnt.deferred = true;
            return;
        }
    }
    constexpr int NT = N > 2 ? N : 2;
    double time[NT] = {};
    int roots = 0;
    // The partition's cut points: the first `roots` of time[], clamped.
    auto set_cuts = [&](int n) {
        out.nb = n;
#pragma unroll
        for (int k = 0; k < NT; k++)
            if (k < n) out.c[k] = clamp_cut(time[k]);
    };
    if (reduced > 2) {
        if (!could_have_roots<N>(op, reduced, pos)) return;
        if constexpr (screens_sextic<CT>::value && N == 6) {
            // RootIsolator<6> starts with the same test and returns no root when it finds no sign change.
            if (reduced != 6 || bernstein_variations<6>(op) != 0) {
                cnt.deferred = true;
                return;
            }
            set_cuts(0);
            check_piece<N>(0, op, reduced, out, pos);
            return;
        }
        cnt(S_ISOLATE);
        if constexpr (stops_at_solver<CT>::value) return;
        roots = isolate_roots<N, NT, plain_roots<CT>::value>(op, reduced, time);
        // Pieces between consecutive roots (all in [0,1], ascending) and the ends, as CTCD.cpp:160-176.
        set_cuts(roots);
#pragma unroll
        for (int p = 0; p <= NT; p++)
            if (p <= roots) check_piece<N>(p, op, reduced, out, pos);
        return;
    } else if (reduced == 2) {
        roots = quad_roots(op[0], op[1], op[2], time[0], time[1]);
        if (roots) cnt(S_QUAD_ROOT);
        // std::sort of two values: swaps only when time[1] < time[0].
        if (roots == 2 && time[1] < time[0]) {
            double s = time[0];
            time[0] = time[1];
            time[1] = s;
        }
    } else if (reduced == 1) {
        time[0] = -op[1] / op[0];
        roots = 1;
    } else {
        if ((!pos && op[0] <= 0) || (pos && op[0] >= 0)) out.keep = 1u;   // the piece [0, 1]
        return;
    }
    set_cuts(roots);
    if (roots > 0) {
        if (time[0] >= 0) check_piece<N>(0, op, reduced, out, pos);
        if (roots == 2) {
            if (!((time[0] < 0 && time[1] < 0) || (time[0] > 1.0 && time[1] > 1.0)))
                check_piece<N>(1, op, reduced, out, pos);
        }
        double last = roots == 2 ? time[1] : time[0];
        if (last <= 1.0) check_piece<N>(roots, op, reduced, out, pos);
    } else {
        check_piece<N>(0, op, reduced, out, pos);
    }
}

// ------------------------------------------------------------------------------------------------
// Coefficient builders (CTCD.cpp:179-257)
// ------------------------------------------------------------------------------------------------
template <int CAP, class Cn>
CCD_HD void barycentric_poly(V3 x10, V3 x20, V3 x30, V3 v10, V3 v20, V3 v30, Ivls<CAP> &out, Cn &cnt)
{
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
    find_intervals<4>(op, out, true, cnt);
}

template <int CAP, class Cn>
CCD_HD void plane_poly(V3 x10, V3 x20, V3 x30, V3 v10, V3 v20, V3 v30, Ivls<CAP> &out, Cn &cnt)
{
    double op[4];
    V3 v23 = cross(v20, v30);
    V3 x2v3 = cross(x20, v30);
    V3 v2x3 = cross(v20, x30);
    V3 x23 = cross(x20, x30);
    op[0] = dot(v10, v23);
    op[1] = dot(x10, v23) + dot(v10, x2v3) + dot(v10, v2x3);
    op[2] = dot(x10, x2v3) + dot(x10, v2x3) + dot(v10, x23);
    op[3] = dot(x10, x23);
    find_intervals<3>(op, out, true, cnt);
}

template <int CAP, class Cn>
CCD_HD void distance_poly(V3 x10, V3 x20, V3 x30, V3 v10, V3 v20, V3 v30, double minD, Ivls<CAP> &out, Cn &cnt)
{
    V3 v23 = cross(v20, v30);
    V3 x2v3 = cross(x20, v30);
    V3 v2x3 = cross(v20, x30);
    V3 x23 = cross(x20, x30);
    double A = dot(v10, v23);
    double B = dot(x10, v23) + dot(v10, x2v3) + dot(v10, v2x3);
    double C = dot(x10, x2v3) + dot(x10, v2x3) + dot(v10, x23);
    double D = dot(x10, x23);
    V3 E = x23;
    V3 F = x2v3 + v2x3;
    V3 G = v23;
    double op[7];
    op[0] = A * A;
    op[1] = 2 * A * B;
    op[2] = B * B + 2 * A * C - dot(G, G) * minD;
    op[3] = 2 * A * D + 2 * B * C - 2 * dot(G, F) * minD;
    op[4] = 2 * B * D + C * C - (2 * dot(G, E) + dot(F, F)) * minD;
    op[5] = 2 * C * D - 2 * dot(F, E) * minD;
    op[6] = D * D - dot(E, E) * minD;
    find_intervals<6>(op, out, false, cnt);
}

// ------------------------------------------------------------------------------------------------
// Primitive tests (CTCD.cpp:259-692).  Each returns true on collision and then sets t.
// ------------------------------------------------------------------------------------------------

// Pruning.  A primitive reports a collision only through a combination of one interval from each of
// its lists whose members overlap pairwise (CTCD.cpp:352-396, :477-507, :575-601), and the lists are
// computed one after another.  When no choice from the lists computed so far overlaps pairwise, no later
// list can produce a combination, so the result is false whatever the later root solves return; the
// product kernels return false there without solving them (the reference solves them and then finds no
// combination).  `mask` selects the first list's members (the coplanar intervals of edge-edge).
// Lists are walked piece by piece over their kept pieces (set bits, ascending: the list order).
// Intervals on a line overlap pairwise exactly when the largest lower end is at most the smallest upper
// end (closed intervals; TimeInterval::overlap tests pairs).  A list's ends are clamped cut points, 0 or
// 1: finite, in [+0, 1], never NaN or -0 (clamp_cut maps both to +0), so the running maximum and
// minimum below decide exactly what the pairwise test decides, and a branch whose running interval is
// already empty is left early.
template <int A, int B>
CCD_HD bool any_overlap(const Ivls<A> &a, unsigned mask, const Ivls<B> &b)
{
    for (unsigned mi = a.keep & mask; mi; mi &= mi - 1) {
        double l0, u0;
        a.get(low_bit(mi), l0, u0);
        for (unsigned mj = b.keep; mj; mj &= mj - 1) {
            double l1, u1;
            b.get(low_bit(mj), l1, u1);
            if (smax(l0, l1) <= smin(u0, u1)) return true;
        }
    }
    return false;
}
template <int A, int B, int C>
CCD_HD bool any_overlap(const Ivls<A> &a, unsigned mask, const Ivls<B> &b, const Ivls<C> &c)
{
    for (unsigned mi = a.keep & mask; mi; mi &= mi - 1) {
        double l0, u0;
        a.get(low_bit(mi), l0, u0);
        for (unsigned mj = b.keep; mj; mj &= mj - 1) {
            double l1, u1;
            b.get(low_bit(mj), l1, u1);
            l1 = smax(l0, l1);
            u1 = smin(u0, u1);
            if (l1 > u1) continue;
            for (unsigned mk = c.keep; mk; mk &= mk - 1) {
                double l2, u2;
                c.get(low_bit(mk), l2, u2);
                if (smax(l1, l2) <= smin(u1, u2)) return true;
            }
        }
    }
    return false;
}
template <int A, int B, int C, int D>
CCD_HD bool any_overlap(const Ivls<A> &a, unsigned mask, const Ivls<B> &b, const Ivls<C> &c, const Ivls<D> &d)
{
    for (unsigned mi = a.keep & mask; mi; mi &= mi - 1) {
        double l0, u0;
        a.get(low_bit(mi), l0, u0);
        for (unsigned mj = b.keep; mj; mj &= mj - 1) {
            double l1, u1;
            b.get(low_bit(mj), l1, u1);
            l1 = smax(l0, l1);
            u1 = smin(u0, u1);
            if (l1 > u1) continue;
            for (unsigned mk = c.keep; mk; mk &= mk - 1) {
                double l2, u2;
                c.get(low_bit(mk), l2, u2);
                l2 = smax(l1, l2);
                u2 = smin(u1, u2);
                if (l2 > u2) continue;
                for (unsigned mm = d.keep; mm; mm &= mm - 1) {
                    double l3, u3;
                    d.get(low_bit(mm), l3, u3);
                    if (smax(l2, l3) <= smin(u2, u3)) return true;
                }
            }
        }
    }
    return false;
}

// edgeEdgeCTCD (CTCD.cpp:259-411): edges (q0,p0) and (q1,p1).
template <class C = NoCount>
CCD_HD bool edge_edge(V3 q0s, V3 p0s, V3 q1s, V3 p1s, V3 q0e, V3 p0e, V3 q1e, V3 p1e, double eta, double &t,
                      C &&cnt = C())
{
    cnt(S_EE_DIST);
    double minD = eta * eta;
    V3 vp0 = p0e - p0s;
    V3 vp1 = p1e - p1s;
    V3 vq0 = q0e - q0s;
    V3 vq1 = q1e - q1s;
    Ivls<7> raw;
    distance_poly(p0s - p1s, p0s - q0s, p1s - q1s, vp0 - vp1, vp0 - vq0, vp1 - vq1, minD, raw, cnt);
    if constexpr (screens_sextic<C>::value) {
        if (raw.keep != 0) cnt.deferred = true;   // coplanar intervals: the full test decides
        return false;
    }

    // Parallel-edge classification of each raw interval at its midpoint (CTCD.cpp:294-308).  The
    // reference copies the raw intervals into a coplanar and a parallel list; here they stay in `raw`
    // and bit p of `par` marks piece p parallel.  Both lists keep the raw order, which is all the
    // combination step below depends on, and two fewer 7-interval lists stay live across the four
    // barycentric solves.
    unsigned par = 0;
    int ncop = 0;
    for (unsigned m = raw.keep; m; m &= m - 1) {
        const int p = low_bit(m);
        cnt(S_EE_PAR);
        double rl, ru;
        raw.get(p, rl, ru);
        double midt = (ru + rl) / 2;
        V3 a = (q0s - p0s) + midt * (vq0 - vp0);
        V3 b = (q1s - p1s) + midt * (vq1 - vp1);
        if (norm(cross(a, b)) < 1e-8)
            par |= 1u << p;
        else
            ncop++;
    }
    if (ncop == 0) return false;

    Ivls<5> a0, a1, b0, b1;
    V3 x10 = p1s - q1s, v10 = vp1 - vq1;
    cnt(S_EE_BARY);
    barycentric_poly(x10, p0s - q0s, q0s - q1s, v10, vp0 - vq0, vq0 - vq1, a0, cnt);
    if (a0.empty()) return false;
    if (prunes<C>::value && !any_overlap(raw, ~par, a0)) return false;
    cnt(S_EE_BARY);
    barycentric_poly(x10, q0s - p0s, p0s - q1s, v10, vq0 - vp0, vp0 - vq1, a1, cnt);
    if (a1.empty()) return false;
    if (prunes<C>::value && !any_overlap(raw, ~par, a0, a1)) return false;
    x10 = p0s - q0s;
    v10 = vp0 - vq0;
    cnt(S_EE_BARY);
    barycentric_poly(x10, p1s - q1s, q1s - q0s, v10, vp1 - vq1, vq1 - vq0, b0, cnt);
    if (b0.empty()) return false;
    if (prunes<C>::value && !any_overlap(raw, ~par, a0, a1, b0)) return false;
    cnt(S_EE_BARY);
    barycentric_poly(x10, q1s - p1s, p1s - q0s, v10, vq1 - vp1, vp1 - vq0, b1, cnt);
    if (b1.empty()) return false;

    // All combinations (CTCD.cpp:352-396): a combination counts if its five intervals overlap
    // pairwise and their intersection touches no parallel interval; t is the smallest lower end.  The
    // intersection [il, iu] is the running maximum of the lower ends and minimum of the upper ends
    // (exact: see any_overlap), and a combination is valid exactly when il <= iu.
    bool col = false;
    double mint = 1.0;
    for (unsigned mi = raw.keep & ~par; mi; mi &= mi - 1) {   // the coplanar intervals, in order
        double l0, u0;
        raw.get(low_bit(mi), l0, u0);
        for (unsigned mj = a0.keep; mj; mj &= mj - 1) {
            double l1, u1;
            a0.get(low_bit(mj), l1, u1);
            l1 = smax(l0, l1);
            u1 = smin(u0, u1);
            if (l1 > u1) continue;
            for (unsigned mk = a1.keep; mk; mk &= mk - 1) {
                double l2, u2;
                a1.get(low_bit(mk), l2, u2);
                l2 = smax(l1, l2);
                u2 = smin(u1, u2);
                if (l2 > u2) continue;
                for (unsigned ml = b0.keep; ml; ml &= ml - 1) {
                    double l3, u3;
                    b0.get(low_bit(ml), l3, u3);
                    l3 = smax(l2, l3);
                    u3 = smin(u2, u3);
                    if (l3 > u3) continue;
                    for (unsigned mm = b1.keep; mm; mm &= mm - 1) {
                        double l4, u4;
                        b1.get(low_bit(mm), l4, u4);
                        const double il = smax(l3, l4), iu = smin(u3, u4);
                        if (il > iu) continue;
                        bool skip = false;
                        for (unsigned mp = par; mp; mp &= mp - 1) {
                            double pl, pu;
                            raw.get(low_bit(mp), pl, pu);
                            if (!(il > pu || pl > iu)) skip = true;
                        }
                        if (!skip) {
                            mint = smin(mint, il);
                            col = true;
                        }
                    }
                }
            }
        }
    }
    if (col) t = mint;
    return col;
}

// vertexFaceCTCD (CTCD.cpp:413-508): vertex q0 and face (q1,q2,q3).
template <class C = NoCount>
CCD_HD bool vertex_face(V3 q0s, V3 q1s, V3 q2s, V3 q3s, V3 q0e, V3 q1e, V3 q2e, V3 q3e, double eta, double &t,
                        C &&cnt = C())
{
    cnt(S_VF_CALL);
    double minD = eta * eta;
    V3 v0 = q0e - q0s;
    V3 v1 = q1e - q1s;
    V3 v2 = q2e - q2s;
    V3 v3 = q3e - q3s;
    Ivls<4> e1, e2, e3;
    cnt(S_VF_PLANE);
    plane_poly(q0s - q1s, cross(q3s - q1s, q2s - q1s), q3s - q1s, v0 - v1, cross(v3 - v1, v2 - v1), v3 - v1, e1, cnt);
    if (e1.empty()) return false;
    cnt(S_VF_PLANE);
    plane_poly(q0s - q2s, cross(q1s - q2s, q3s - q2s), q1s - q2s, v0 - v2, cross(v1 - v2, v3 - v2), v1 - v2, e2, cnt);
    if (e2.empty()) return false;
    if (prunes<C>::value && !any_overlap(e1, ~0u, e2)) return false;
    cnt(S_VF_PLANE);
    plane_poly(q0s - q3s, cross(q2s - q3s, q1s - q3s), q2s - q3s, v0 - v3, cross(v2 - v3, v1 - v3), v2 - v3, e3, cnt);
    if (e3.empty()) return false;
    if (prunes<C>::value && !any_overlap(e1, ~0u, e2, e3)) return false;
    Ivls<7> cop;
    cnt(S_VF_DIST);
    distance_poly(q0s - q1s, q2s - q1s, q3s - q1s, v0 - v1, v2 - v1, v3 - v1, minD, cop, cnt);
    if (cop.empty()) return false;

    bool col = false;
    double mint = 1.0;
    for (unsigned mi = cop.keep; mi; mi &= mi - 1) {
        double L[4], U[4];
        cop.get(low_bit(mi), L[0], U[0]);
        for (unsigned mj = e1.keep; mj; mj &= mj - 1) {
            e1.get(low_bit(mj), L[1], U[1]);
            for (unsigned mk = e2.keep; mk; mk &= mk - 1) {
                e2.get(low_bit(mk), L[2], U[2]);
                for (unsigned ml = e3.keep; ml; ml &= ml - 1) {
                    e3.get(low_bit(ml), L[3], U[3]);
                    bool ov = true;
#pragma unroll
                    for (int x = 0; x < 4; x++)
#pragma unroll
                        for (int y = x + 1; y < 4; y++)
                            if (L[x] > U[y] || L[y] > U[x]) ov = false;
                    if (!ov) continue;
                    double il = 0.0;
#pragma unroll
                    for (int x = 0; x < 4; x++) il = smax(L[x], il);
                    mint = smin(il, mint);
                    col = true;
                }
            }
        }
    }
    if (col) t = mint;
    return col;
}

// vertexEdgeCTCD's two side quadratics (CTCD.cpp:520-555): when the vertex projects between the edge's
// ends.  False when either interval list is empty (the reference returns there); e2 is computed only
// if e1 is not empty, as in the reference.
template <class Cn>
CCD_HD bool ve_sides(V3 q0s, V3 q1s, V3 q2s, V3 q0e, V3 q1e, V3 q2e, Ivls<3> &e1, Ivls<3> &e2, Cn &cnt)
{
    V3 v0 = q0e - q0s;
    V3 v1 = q1e - q1s;
    V3 v2 = q2e - q2s;
    V3 ab = q2s - q1s;
    V3 ac = q0s - q1s;
    V3 cb = q2s - q0s;
    V3 vab = v2 - v1;
    V3 vac = v0 - v1;
    V3 vcb = v2 - v0;
    double op[3];
    op[0] = dot(vab, vac);
    op[1] = dot(ac, vab) + dot(ab, vac);
    op[2] = dot(ab, ac);
    cnt(S_VE_QUAD);
    find_intervals<2>(op, e1, true, cnt);
    if (e1.empty()) return false;
    op[0] = dot(vab, vcb);
    op[1] = dot(cb, vab) + dot(ab, vcb);
    op[2] = dot(ab, cb);
    cnt(S_VE_QUAD);
    find_intervals<2>(op, e2, true, cnt);
    return !e2.empty();
}

// Which vertex-edge sub-tests of a stencil's window would reach their quartic (both side quadratics
// non-empty and overlapping): bit j for the j-th vertex-edge test in the reference's order.  Cheap
// (closed-form quadratics only); the narrowphase groups stencils by it so that a wavefront's lanes
// take the same quartic solves.
struct SolverProbe {
    static constexpr bool stops_at_solver = true;
    bool called = false;
    CCD_FHD void operator()(int stage)
    {
        if (stage == S_ISOLATE) called = true;
    }
};

// Which of the polynomials of a window's face or edge-edge test would reach the root solver, each
// evaluated whether or not the test would get that far (vertexFaceCTCD's / edgeEdgeCTCD's arguments for
// check_vf_window's / check_ee_window's vertex order).  Vertex-face: the three edge-side cubics (bits
// 0-2) and the coplanar-distance sextic (bit 3).  Edge-edge: the coplanar-distance sextic (bit 0) and
// how many of the four barycentric quartics (bits 1-3).  A sort key only.
CCD_HD unsigned first_phase_signature(bool vf, const V3 *s, const V3 *e, double eta)
{
    const V3 v0 = e[0] - s[0], v1 = e[1] - s[1], v2 = e[2] - s[2], v3 = e[3] - s[3];
    if (vf) {
        Ivls<4> iv;
        Ivls<7> iv7;
        SolverProbe p0, p1, p2, p3;
        plane_poly(s[0] - s[1], cross(s[3] - s[1], s[2] - s[1]), s[3] - s[1], v0 - v1, cross(v3 - v1, v2 - v1), v3 - v1,
                   iv, p0);
        plane_poly(s[0] - s[2], cross(s[1] - s[2], s[3] - s[2]), s[1] - s[2], v0 - v2, cross(v1 - v2, v3 - v2), v1 - v2,
                   iv, p1);
        plane_poly(s[0] - s[3], cross(s[2] - s[3], s[1] - s[3]), s[2] - s[3], v0 - v3, cross(v2 - v3, v1 - v3), v2 - v3,
                   iv, p2);
        distance_poly(s[0] - s[1], s[2] - s[1], s[3] - s[1], v0 - v1, v2 - v1, v3 - v1, eta * eta, iv7, p3);
        return (p0.called ? 1u : 0u) | (p1.called ? 2u : 0u) | (p2.called ? 4u : 0u) | (p3.called ? 8u : 0u);
    }
    // edgeEdgeCTCD(q0 = s[0], p0 = s[1], q1 = s[2], p1 = s[3])
    Ivls<7> iv7;
    Ivls<5> iv5;
    SolverProbe pd, pa0, pa1, pb0, pb1;
    distance_poly(s[1] - s[3], s[1] - s[0], s[3] - s[2], v1 - v3, v1 - v0, v3 - v2, eta * eta, iv7, pd);
    barycentric_poly(s[3] - s[2], s[1] - s[0], s[0] - s[2], v3 - v2, v1 - v0, v0 - v2, iv5, pa0);
    barycentric_poly(s[3] - s[2], s[0] - s[1], s[1] - s[2], v3 - v2, v0 - v1, v1 - v2, iv5, pa1);
    barycentric_poly(s[1] - s[0], s[3] - s[2], s[2] - s[0], v1 - v0, v3 - v2, v2 - v0, iv5, pb0);
    barycentric_poly(s[1] - s[0], s[2] - s[3], s[3] - s[0], v1 - v0, v2 - v3, v3 - v0, iv5, pb1);
    const unsigned nb = (pa0.called ? 1u : 0u) + (pa1.called ? 1u : 0u) + (pb0.called ? 1u : 0u) + (pb1.called ? 1u : 0u);
    return (pd.called ? 1u : 0u) | (nb << 1);
}

CCD_HD unsigned ve_signature(bool vf, const V3 *s, const V3 *e)
{
    unsigned sig = 0;
    NoCount nc;
    auto bit = [&](int j, int a, int b, int c) {
        Ivls<3> e1, e2;
        if (ve_sides(s[a], s[b], s[c], e[a], e[b], e[c], e1, e2, nc) && any_overlap(e1, ~0u, e2)) sig |= 1u << j;
    };
    if (vf) {   // check_vf_window's order
        bit(0, 0, 1, 2);
        bit(1, 0, 2, 3);
        bit(2, 0, 3, 1);
    } else {    // check_ee_window's order
        bit(0, 0, 2, 3);
        bit(1, 1, 2, 3);
        bit(2, 2, 0, 1);
        bit(3, 3, 0, 1);
    }
    return sig;
}

// vertexEdgeCTCD (CTCD.cpp:511-602): vertex q0 and edge (q1,q2).
template <class Cn = NoCount>
CCD_HD bool vertex_edge(V3 q0s, V3 q1s, V3 q2s, V3 q0e, V3 q1e, V3 q2e, double eta, double &t, Cn &&cnt = Cn())
{
    cnt(S_VE_CALL);
    double minD = eta * eta;
    V3 v0 = q0e - q0s;
    V3 v1 = q1e - q1s;
    V3 v2 = q2e - q2s;
    V3 ab = q2s - q1s;
    V3 ac = q0s - q1s;
    V3 vab = v2 - v1;
    V3 vac = v0 - v1;
    Ivls<3> e1, e2;
    Ivls<5> colin;
    if (!ve_sides(q0s, q1s, q2s, q0e, q1e, q2e, e1, e2, cnt)) return false;
    if (prunes<Cn>::value && !any_overlap(e1, ~0u, e2)) return false;
    double op[5];
    cnt(S_VE_QUART);
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
    find_intervals<4>(op, colin, false, cnt);
    if (colin.empty()) return false;

    bool col = false;
    double mint = 1.0;
    for (unsigned mi = colin.keep; mi; mi &= mi - 1) {
        double L[3], U[3];
        colin.get(low_bit(mi), L[0], U[0]);
        for (unsigned mj = e1.keep; mj; mj &= mj - 1) {
            e1.get(low_bit(mj), L[1], U[1]);
            for (unsigned mk = e2.keep; mk; mk &= mk - 1) {
                e2.get(low_bit(mk), L[2], U[2]);
                bool ov = true;
#pragma unroll
                for (int x = 0; x < 3; x++)
#pragma unroll
                    for (int y = x + 1; y < 3; y++)
                        if (L[x] > U[y] || L[y] > U[x]) ov = false;
                if (!ov) continue;
                double il = 0.0;
#pragma unroll
                for (int x = 0; x < 3; x++) il = smax(L[x], il);
                mint = smin(il, mint);
                col = true;
            }
        }
    }
    if (col) t = mint;
    return col;
}

// vertexVertexCTCD (CTCD.cpp:604-692)
template <class Cn = NoCount>
CCD_HD bool vertex_vertex(V3 q1s, V3 q2s, V3 q1e, V3 q2e, double eta, double &t, Cn &&cnt = Cn())
{
    cnt(S_VV);
    double min_d = eta * eta;
    double t1 = 0, t2 = 0;
    int roots = 0;
    V3 v1 = q1e - q1s;
    V3 v2 = q2e - q2s;
    double a = dot(v1, v1) + dot(v2, v2) - 2 * dot(v1, v2);
    double b = 2 * (dot(v1, q1s) - dot(v2, q1s) - dot(v1, q2s) + dot(v2, q2s));
    double c = dot(q1s, q1s) + dot(q2s, q2s) - 2 * dot(q1s, q2s) - min_d;
    if (a != 0) {
        roots = quad_roots(a, b, c, t1, t2);
        if (roots) cnt(S_QUAD_ROOT);
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
    if (roots == 2) {
        if (check_interval_passes<2>(0, t1, op, 2, false)) {
            t = 0;
            return true;
        }
        if (check_interval_passes<2>(t1, t2, op, 2, false)) {
            t = t1;
            return true;
        }
        if (check_interval_passes<2>(t2, 1.0, op, 2, false)) {
            t = t2;
            return true;
        }
        return false;
    } else if (roots == 1) {
        if (check_interval_passes<2>(0, t1, op, 2, false)) {
            t = 0;
            return true;
        }
        if (check_interval_passes<2>(t1, 1.0, op, 2, false)) {
            t = t1;
            return true;
        }
        return false;
    }
    if (check_interval_passes<2>(0, 1.0, op, 2, false)) {
        t = 0;
        return true;
    }
    return false;
}

// ------------------------------------------------------------------------------------------------
// Stencil narrowphase for one stitched window (CTCDNarrowPhase.cpp:43-69 and :91-132).
// p[0..3] = start positions, e[0..3] = end positions of the stencil's vertices in key order.
// Returns the firing sub-test (0 = VF/EE, then VE in reference order, then VV) or -1.
// ------------------------------------------------------------------------------------------------
template <class C = NoCount>
CCD_HD int check_vf_window(const V3 *s, const V3 *e, double eta, double &t, C &&cnt = C())
{
    if (vertex_face(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, t, cnt)) return 0;
    if (vertex_edge(s[0], s[1], s[2], e[0], e[1], e[2], eta, t, cnt)) return 1;
    if (vertex_edge(s[0], s[2], s[3], e[0], e[2], e[3], eta, t, cnt)) return 2;
    if (vertex_edge(s[0], s[3], s[1], e[0], e[3], e[1], eta, t, cnt)) return 3;
    if (vertex_vertex(s[0], s[1], e[0], e[1], eta, t, cnt)) return 4;
    if (vertex_vertex(s[0], s[2], e[0], e[2], eta, t, cnt)) return 5;
    if (vertex_vertex(s[0], s[3], e[0], e[3], eta, t, cnt)) return 6;
    return -1;
}

template <class C = NoCount>
CCD_HD int check_ee_window(const V3 *s, const V3 *e, double eta, double &t, C &&cnt = C())
{
    if (edge_edge(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, t, cnt)) return 0;
    if (vertex_edge(s[0], s[2], s[3], e[0], e[2], e[3], eta, t, cnt)) return 1;
    if (vertex_edge(s[1], s[2], s[3], e[1], e[2], e[3], eta, t, cnt)) return 2;
    if (vertex_edge(s[2], s[0], s[1], e[2], e[0], e[1], eta, t, cnt)) return 3;
    if (vertex_edge(s[3], s[0], s[1], e[3], e[0], e[1], eta, t, cnt)) return 4;
    if (vertex_vertex(s[0], s[2], e[0], e[2], eta, t, cnt)) return 5;
    if (vertex_vertex(s[0], s[3], e[0], e[3], eta, t, cnt)) return 6;
    if (vertex_vertex(s[1], s[2], e[1], e[2], eta, t, cnt)) return 7;
    if (vertex_vertex(s[1], s[3], e[1], e[3], eta, t, cnt)) return 8;
    return -1;
}

// ------------------------------------------------------------------------------------------------
// Proximity tests (include/Distance.h:14-174), straight-line FP64 in the reference's expression order.
// ------------------------------------------------------------------------------------------------

// Distance.h:14-19: is p closer than eta to the plane of (q0, q1, q2)?  No division.
CCD_HD bool vertex_plane_distance_less_than(V3 p, V3 q0, V3 q1, V3 q2, double eta)
{
    V3 c = cross(q1 - q0, q2 - q0);
    return dot(c, p - q0) * dot(c, p - q0) < eta * eta * dot(c, c);
}

// Distance.h:23-28: is the line (p0,p1) closer than eta to the line (q0,q1)?  No division.
CCD_HD bool line_line_distance_less_than(V3 p0, V3 p1, V3 q0, V3 q1, double eta)
{
    V3 c = cross(p1 - p0, q1 - q0);
    return dot(c, q0 - p0) * dot(c, q0 - p0) < eta * eta * dot(c, c);
}

CCD_FHD double clamp01(double u) { return smin(1.0, smax(u, 0.0)); }   // Distance.h:183-186

// Distance.h:32-115: vector from p to its closest point on the triangle (q0,q1,q2), by Voronoi region,
// with the closest point's barycentric coordinates.
CCD_HD V3 vertex_face_distance(V3 p, V3 q0, V3 q1, V3 q2, double &b0, double &b1, double &b2)
{
    V3 ab = q1 - q0;
    V3 ac = q2 - q0;
    V3 ap = p - q0;
    double d1 = dot(ab, ap);
    double d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) {
        b0 = 1.0;
        b1 = 0.0;
        b2 = 0.0;
        return q0 - p;
    }
    V3 bp = p - q1;
    double d3 = dot(ab, bp);
    double d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
        b0 = 0.0;
        b1 = 1.0;
        b2 = 0.0;
        return q1 - p;
    }
    double vc = d1 * d4 - d3 * d2;
    if ((vc <= 0) && (d1 >= 0) && (d3 <= 0)) {
        double v = d1 / (d1 - d3);
        b0 = 1.0 - v;
        b1 = v;
        b2 = 0;
        return (q0 + v * ab) - p;
    }
    V3 cp = p - q2;
    double d5 = dot(ab, cp);
    double d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) {
        b0 = 0;
        b1 = 0;
        b2 = 1.0;
        return q2 - p;
    }
    double vb = d5 * d2 - d1 * d6;
    if ((vb <= 0) && (d2 >= 0) && (d6 <= 0)) {
        double w = d2 / (d2 - d6);
        b0 = 1 - w;
        b1 = 0;
        b2 = w;
        return (q0 + w * ac) - p;
    }
    double va = d3 * d6 - d5 * d4;
    if ((va <= 0) && (d4 - d3 >= 0) && (d5 - d6 >= 0)) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        b0 = 0;
        b1 = 1.0 - w;
        b2 = w;
        return (q1 + w * (q2 - q1)) - p;
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom;
    double w = vc * denom;
    double u = 1.0 - v - w;
    b0 = u;
    b1 = v;
    b2 = w;
    return ((u * q0 + v * q1) + w * q2) - p;
}

// Distance.h:119-174: shortest vector from segment (p0,p1) to segment (q0,q1), with the closest
// points' barycentric coordinates (parallel or degenerate segments handled as the reference does).
CCD_HD V3 edge_edge_distance(V3 p0, V3 p1, V3 q0, V3 q1, double &p0b, double &p1b, double &q0b, double &q1b)
{
    V3 d1 = p1 - p0;
    V3 d2 = q1 - q0;
    V3 r = p0 - q0;
    double a = dot(d1, d1);
    double e = dot(d2, d2);
    double f = dot(d2, r);
    double s, t;
    double c = dot(d1, r);
    double b = dot(d1, d2);
    double denom = a * e - b * b;
    if (denom != 0.0)
        s = clamp01((b * f - c * e) / denom);
    else
        s = 0;
    double tnom = b * s + f;
    if (tnom < 0 || e == 0) {
        t = 0;
        s = (a == 0) ? 0 : clamp01(-c / a);
    } else if (tnom > e) {
        t = 1.0;
        s = (a == 0) ? 0 : clamp01((b - c) / a);
    } else {
        t = tnom / e;
    }
    V3 c1 = p0 + s * d1;
    V3 c2 = q0 + t * d2;
    p0b = 1.0 - s;
    p1b = s;
    q0b = 1.0 - t;
    q1b = t;
    return c2 - c1;
}

// The position History::stitchCommonHistory assigns at t = 0 for a two-entry history
// (History.cpp:127-133 with dt = 1, a = 0, b = 1): (1*start + 0*end) / 1, kept literally so that
// signed zeros come out as in the reference.
CCD_FHD double stitch0(double s, double e) { return (1.0 * s + 0.0 * e) / 1.0; }

// ------------------------------------------------------------------------------------------------
// Separating-plane narrowphase (SeparatingPlaneNarrowPhase.cpp:9-278) over a CSR history: vertex v
// owns entries off[v] .. off[v+1]-1 with times t[] (first 0, last 1, nondecreasing) and positions
// p[3*e ..].  Every expression follows the reference's order (Eigen dot as dot() above).
// ------------------------------------------------------------------------------------------------
struct HistCSR {
    const int64_t *off;
    const double *t;
    const double *p;
};
CCD_FHD V3 hist_pos(const HistCSR &h, int64_t e) { return mk(h.p[3 * e], h.p[3 * e + 1], h.p[3 * e + 2]); }

// History::getPosAtTime (History.cpp:47-76).  Its bisection only narrows the start of a forward scan
// that stops at the first entry later than `time`, so the result is that of an upper-bound search:
// next = first entry with time > `time`; idx = next - 1 (relative to the vertex's first entry), or the
// last entry when none is later.  Times passed here lie in [0, 1].
CCD_HD V3 pos_at_time(const HistCSR &h, int v, double time, int &idx)
{
    const int64_t b = h.off[v], e = h.off[v + 1];
    int64_t lo = b, hi = e;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (h.t[mid] <= time)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo == e) {
        idx = (int)(e - 1 - b);
        return hist_pos(h, e - 1);
    }
    const int64_t prev = lo > b ? lo - 1 : b;
    idx = (int)(prev - b);
    double dt = h.t[lo] - h.t[prev];
    double alpha = (time - h.t[prev]) / dt;
    V3 a = hist_pos(h, prev), c = hist_pos(h, lo);
    return (1.0 - alpha) * a + alpha * c;
}

// planeIntersect (SeparatingPlaneNarrowPhase.cpp:267-278)
CCD_FHD double plane_intersect(V3 ppos, V3 pvel, V3 nrm, V3 old_p, V3 new_p, double dt, double eta)
{
    double numerator = 0.5 * eta - dot(old_p - ppos, nrm);
    double denom = -dot(pvel, nrm);
    if (dt != 0.0) denom += dot(new_p - old_p, nrm) / dt;
    return numerator / denom;
}

// planeTrajectoryIntersect (SeparatingPlaneNarrowPhase.cpp:216-265); start is relative to the vertex's
// first entry.  Returns the time from tstart until the vertex trajectory meets the moving plane, or
// +infinity.
CCD_HD double plane_trajectory_intersect(const HistCSR &h, int v, int start, bool forward, V3 ppos, V3 pvel, V3 nrm,
                                         V3 ptstart, double tstart, double eta)
{
    const double inf = bits_to_double(0x7ff0000000000000ll);
    const int64_t b = h.off[v];
    const int n = (int)(h.off[v + 1] - b);
    double dt = forward ? h.t[b + start] - tstart : tstart - h.t[b + start];
    double t = plane_intersect(ppos, (forward ? 1.0 : -1.0) * pvel, nrm, ptstart, hist_pos(h, b + start), dt, eta);
    if (t >= 0 && t <= dt) return t;
    if (forward) {
        for (int i = start; i < n; i++) {
            if (i == n - 1) return inf;
            V3 npp = ppos + (h.t[b + i] - tstart) * pvel;
            double d = h.t[b + i + 1] - h.t[b + i];
            double u = plane_intersect(npp, pvel, nrm, hist_pos(h, b + i), hist_pos(h, b + i + 1), d, eta);
            if (u >= 0 && u <= d) return h.t[b + i] + u - tstart;
        }
    } else {
        for (int i = start; i >= 0; i--) {
            if (i == 0) return inf;
            V3 npp = ppos - (tstart - h.t[b + i]) * pvel;
            double d = h.t[b + i] - h.t[b + i - 1];
            double u = plane_intersect(npp, -1.0 * pvel, nrm, hist_pos(h, b + i), hist_pos(h, b + i - 1), d, eta);
            if (u >= 0 && u <= d) return tstart - (h.t[b + i] - u);
        }
    }
    return inf;
}

// checkInterval (SeparatingPlaneNarrowPhase.cpp:63-214) for one stencil (v = the key's vertices), from
// [0, 1].  The reference recurses into the lower sub-interval, then the upper; here a stack of pending
// intervals visits them in the same depth-first order (each node's upper bound is computed before its
// lower subtree is explored; it depends only on the node).  Returns 1 on collision, 0 if none, and 2
// (undecided) where the reference's recursion would not end: a sub-interval equal to its parent (the
// recursion is a pure function of the interval) or one more than SP_MAX_DEPTH levels down (each level
// at most halves the interval; only an interval closing in on a history entry time gets that deep).
constexpr int SP_MAX_DEPTH = 160;

// The trajectory scan one lane runs by itself.  A scanner with the same call operator may split the
// scan over the lanes of a wavefront (every lane then runs the same check on the same stencil).
struct SeqScan {
    CCD_HD double operator()(const HistCSR &h, int v, int start, bool forward, V3 ppos, V3 pvel, V3 nrm, V3 ptstart,
                             double tstart, double eta) const
    {
        return plane_trajectory_intersect(h, v, start, forward, ppos, pvel, nrm, ptstart, tstart, eta);
    }
};

// max_nodes bounds the intervals examined: when more remain the result is 3 ("not decided yet"), which
// a caller uses to sort stencils by how much work they need before running them in full.
template <bool VF, class C = NoCount, class Scan = SeqScan>
CCD_HD int separating_plane_check(const HistCSR &h, const int (&v)[4], double eta, double eps, C &&cnt = C(),
                                  Scan scan = Scan(), int max_nodes = 0x7fffffff)
{
    const double inf = bits_to_double(0x7ff0000000000000ll);
    // pending intervals and their depths; a node at depth d leaves at most d pending uppers below it
    double st_lo[SP_MAX_DEPTH + 2], st_hi[SP_MAX_DEPTH + 2];
    int st_d[SP_MAX_DEPTH + 2];
    st_lo[0] = 0.0;
    st_hi[0] = 1.0;
    st_d[0] = 0;
    int top = 1;
    while (top > 0) {
        if (max_nodes-- <= 0) return 3;
        top--;
        CCD_INDEX_CHECK(top, SP_MAX_DEPTH + 2);
        const double mint = st_lo[top], maxt = st_hi[top];
        const int depth = st_d[top];
        if (depth > SP_MAX_DEPTH) return 2;
        if (maxt < mint) continue;
        if (maxt - mint < 3 * eps) {
            V3 s[4], e[4];
            bool ok = true;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int io = 0, in = 0;
                s[j] = pos_at_time(h, v[j], mint, io);
                e[j] = pos_at_time(h, v[j], maxt, in);
                if (io != in) ok = false;
            }
            if (ok) {
                double t = 0.0;
                int r = VF ? check_vf_window(s, e, eta, t, cnt) : check_ee_window(s, e, eta, t, cnt);
                if (r >= 0) return 1;
                continue;
            }
        }
        const double midt = 0.5 * (mint + maxt);
        V3 mp[4];
        int mi[4];
#pragma unroll
        for (int j = 0; j < 4; j++) mp[j] = pos_at_time(h, v[j], midt, mi[j]);
        double b1 = 0, b2 = 0, b3 = 0, b4 = 0;
        V3 closest = VF ? vertex_face_distance(mp[0], mp[1], mp[2], mp[3], b1, b2, b3)
                        : edge_edge_distance(mp[0], mp[1], mp[2], mp[3], b1, b2, b3, b4);
        if (dot(closest, closest) < eta * eta) return 1;
        // The separating plane: through the midpoint of the closest points, moving with their mean
        // velocity over the current history windows (:162-185).
        V3 dv[4];
        double dts[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t k = h.off[v[j]] + mi[j];
            V3 d = hist_pos(h, k + 1) - hist_pos(h, k);
            dts[j] = h.t[k + 1] - h.t[k];
            dv[j] = d;
        }
        V3 ppos, pvel;
        if (VF) {
            ppos = 0.5 * (((mp[0] + b1 * mp[1]) + b2 * mp[2]) + b3 * mp[3]);
            V3 w0 = mk(dv[0].x / dts[0], dv[0].y / dts[0], dv[0].z / dts[0]);
            V3 w1 = b1 * dv[1], w2 = b2 * dv[2], w3 = b3 * dv[3];
            w1 = mk(w1.x / dts[1], w1.y / dts[1], w1.z / dts[1]);
            w2 = mk(w2.x / dts[2], w2.y / dts[2], w2.z / dts[2]);
            w3 = mk(w3.x / dts[3], w3.y / dts[3], w3.z / dts[3]);
            pvel = 0.5 * (((w0 + w1) + w2) + w3);
        } else {
            ppos = 0.5 * (((b1 * mp[0] + b2 * mp[1]) + b3 * mp[2]) + b4 * mp[3]);
            V3 w0 = b1 * dv[0], w1 = b2 * dv[1], w2 = b3 * dv[2], w3 = b4 * dv[3];
            w0 = mk(w0.x / dts[0], w0.y / dts[0], w0.z / dts[0]);
            w1 = mk(w1.x / dts[1], w1.y / dts[1], w1.z / dts[1]);
            w2 = mk(w2.x / dts[2], w2.y / dts[2], w2.z / dts[2]);
            w3 = mk(w3.x / dts[3], w3.y / dts[3], w3.z / dts[3]);
            pvel = 0.5 * (((w0 + w1) + w2) + w3);
        }
        const double cn = norm(closest);
        closest = mk(closest.x / cn, closest.y / cn, closest.z / cn);
        double t_lo = inf, t_hi = inf;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double sign = (j == 0 || (j == 1 && !VF)) ? -1.0 : 1.0;
            V3 nrm = sign * closest;
            t_lo = smin(t_lo, scan(h, v[j], mi[j], false, ppos, pvel, nrm, mp[j], midt, eta));
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double sign = (j == 0 || (j == 1 && !VF)) ? -1.0 : 1.0;
            V3 nrm = sign * closest;
            t_hi = smin(t_hi, scan(h, v[j], mi[j] + 1, true, ppos, pvel, nrm, mp[j], midt, eta));
        }
        // Depth first, lower sub-interval first: push the upper one underneath.  A child equal to its
        // node is undecided when it is reached (the upper one only if the lower subtree finds nothing).
        if (t_hi < 1.0) {
            const double lo = midt + smax(0.0, t_hi - eps);
            CCD_INDEX_CHECK(top, SP_MAX_DEPTH + 2);
            st_lo[top] = lo;
            st_hi[top] = maxt;
            st_d[top] = lo == mint ? SP_MAX_DEPTH + 1 : depth + 1;   // equal to this node: undecided if reached
            top++;
        }
        if (t_lo < 1.0) {
            const double hi = midt - smax(0.0, t_lo - eps);
            if (hi == maxt) return 2;
            CCD_INDEX_CHECK(top, SP_MAX_DEPTH + 2);
            st_lo[top] = mint;
            st_hi[top] = hi;
            st_d[top] = depth + 1;
            top++;
        }
    }
    return 0;
}

}   // namespace ccd
