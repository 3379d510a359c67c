// Reference-root mode: the reference's polynomial root finder, restated for host and device.
//
// The reference splits [0,1] at the roots that the Jenkins-Traub three-stage algorithm (TOMS 493, in
// C. Bond's C translation, src/rpoly.h:49-759) returns.  On the near-double roots of the
// distance-threshold sextics that algorithm is off by up to ~1e-4, so any other root finder -- however
// exact -- gives times of impact that differ from the reference's by more than 1e-9 on a sizeable
// fraction of hits.  This header reproduces the reference's root finder operation for operation
// (same expressions, same association, same branch structure; FP64, no contraction), so that the
// narrowphase can be run in a mode whose outputs are the reference's bit for bit.
//
// The only library calls of the reference's root finder are sqrt (correctly rounded everywhere), the
// cos/sin of the fixed 94-degree shift rotation (constants: the correctly rounded values, which are
// what the host C library returns), and log/exp in the scaling factor and the root-modulus bound.
// For the latter two this header computes correctly rounded results in double-double arithmetic, so
// host and device agree bit for bit with each other and with any correctly rounded C library.
#pragma once

namespace ccd {
namespace jt {

// ---- double-double helpers (error-free transformations) ----------------------------------------
struct DD {
    double hi, lo;
};

CCD_FHD DD two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
CCD_FHD DD quick_two_sum(double a, double b)
{
    double s = a + b;
    return {s, b - (s - a)};
}
CCD_FHD DD two_prod(double a, double b)
{
    double p = a * b;
    return {p, dfma(a, b, -p)};
}
CCD_FHD DD dd_add(DD a, DD b)
{
    DD s = two_sum(a.hi, b.hi);
    DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
CCD_FHD DD dd_mul(DD a, DD b)
{
    DD p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}
CCD_FHD DD dd_mul_d(DD a, double b)
{
    DD p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return quick_two_sum(p.hi, p.lo);
}
CCD_FHD DD dd_div(DD a, DD b)
{
    double q1 = a.hi / b.hi;
    DD r = dd_add(a, dd_mul_d(b, -q1));
    double q2 = r.hi / b.hi;
    r = dd_add(r, dd_mul_d(b, -q2));
    double q3 = r.hi / b.hi;
    DD q = quick_two_sum(q1, q2);
    return dd_add(q, DD{q3, 0.0});
}

// ln 2 to 106 bits.
constexpr double LN2_HI = 0x1.62e42fefa39efp-1;
constexpr double LN2_LO = 0x1.abc9e3b39803fp-56;

// 1/(2k+1), k = 0..24, and 1/j!, j = 0..27, as double-double (hi, lo).
constexpr double INV_ODD[25][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.5555555555555p-2, 0x1.5555555555555p-56},
    {0x1.999999999999ap-3, -0x1.999999999999ap-57},
    {0x1.2492492492492p-3, 0x1.2492492492492p-57},
    {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
    {0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59},
    {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},
    {0x1.1111111111111p-4, 0x1.1111111111111p-60},
    {0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61},
    {0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59},
    {0x1.8618618618618p-5, 0x1.8618618618618p-59},
    {0x1.642c8590b2164p-5, 0x1.642c8590b2164p-60},
    {0x1.47ae147ae147bp-5, -0x1.eb851eb851eb8p-61},
    {0x1.2f684bda12f68p-5, 0x1.2f684bda12f68p-59},
    {0x1.1a7b9611a7b96p-5, 0x1.1a7b9611a7b96p-61},
    {0x1.0842108421084p-5, 0x1.0842108421084p-60},
    {0x1.f07c1f07c1f08p-6, -0x1.f07c1f07c1f08p-61},
    {0x1.d41d41d41d41dp-6, 0x1.0750750750750p-60},
    {0x1.bacf914c1bad0p-6, -0x1.bacf914c1bad0p-60},
    {0x1.a41a41a41a41ap-6, 0x1.0690690690690p-60},
    {0x1.8f9c18f9c18fap-6, -0x1.f3831f3831f38p-61},
    {0x1.7d05f417d05f4p-6, 0x1.7d05f417d05f4p-62},
    {0x1.6c16c16c16c17p-6, -0x1.f49f49f49f49fp-61},
    {0x1.5c9882b931057p-6, 0x1.310572620ae4cp-61},
    {0x1.4e5e0a72f0539p-6, 0x1.e0a72f0539783p-60},
};
constexpr double INV_FACT[28][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.0000000000000p-1, 0x0.0p+0},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57},
    {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73},
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80},
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
    {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92},
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97},
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
    {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107},
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112},
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
    {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124},
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130},
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},
    {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143},
    {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149},
};

CCD_FHD double bits_ldexp(double x, int e)   // x * 2^e for results in the normal range
{
    while (e > 1000) {
        x *= 0x1p1000;
        e -= 1000;
    }
    while (e < -1000) {
        x *= 0x1p-1000;
        e += 1000;
    }
    return x * bits_to_double((long long)(1023 + e) << 52);
}

// Correctly rounded natural logarithm of a positive finite double (full double-double evaluation: the
// fallback of log_cr when its fast path cannot settle the rounding).
static __attribute__((noinline)) CCD_HD_NI double log_cr_full(double x)
{
    if (!(x > 0.0) || x != x) return x == 0.0 ? -bits_to_double(0x7ff0000000000000ll) : (x - x) / 0.0;
    if (x == bits_to_double(0x7ff0000000000000ll)) return x;
    int e = 0;
    if (x < 0x1p-1000) {   // subnormal or tiny: scale up exactly
        x *= 0x1p200;
        e -= 200;
    }
    long long bits = double_to_bits(x);
    e += (int)((bits >> 52) & 0x7ff) - 1023;
    double m = bits_to_double((bits & 0x000fffffffffffffll) | 0x3ff0000000000000ll);   // [1,2)
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    // log m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716.
    DD num = {m - 1.0, 0.0};   // exact (Sterbenz)
    DD den = two_sum(m, 1.0);
    DD s = dd_div(num, den);
    DD s2 = dd_mul(s, s);
    // sum_{k=0}^{K} s^{2k} / (2k+1), Horner from the top; K = 24 gives |s|^50 < 1e-38.
    DD acc = {INV_ODD[24][0], INV_ODD[24][1]};
#pragma unroll
    for (int k = 23; k >= 0; k--) acc = dd_add(dd_mul(acc, s2), DD{INV_ODD[k][0], INV_ODD[k][1]});
    DD lm = dd_mul(dd_mul_d(s, 2.0), acc);
    DD le = dd_add(two_prod((double)e, LN2_HI), DD{(double)e * LN2_LO, 0.0});
    DD r = dd_add(le, lm);
    return r.hi + r.lo;
}

// Correctly rounded exp of a double whose result is a normal double (full double-double evaluation: the
// fallback of exp_cr when its fast path cannot settle the rounding).
static __attribute__((noinline)) CCD_HD_NI double exp_cr_full(double y)
{
    if (y != y) return y;
    if (y > 709.78) return bits_to_double(0x7ff0000000000000ll);
    if (y < -708.0) return 0.0;   // not reached by the root-modulus bound of a finite polynomial
    double kd = y / LN2_HI;
    kd = kd < 0 ? -(double)(long long)(-kd + 0.5) : (double)(long long)(kd + 0.5);
    int k = (int)kd;
    // r = y - k ln2 in double-double.
    DD r = dd_add(DD{y, 0.0}, dd_mul_d(DD{LN2_HI, LN2_LO}, -kd));
    // exp(r) = sum r^j / j!, |r| <= 0.35: 27 terms reach 1e-40.
    DD acc = {INV_FACT[27][0], INV_FACT[27][1]};
#pragma unroll
    for (int j = 26; j >= 0; j--) acc = dd_add(dd_mul(acc, r), DD{INV_FACT[j][0], INV_FACT[j][1]});
    double v = acc.hi + acc.lo;
    return bits_ldexp(v, k);
}

// ---- fast paths with a rounding test (Ziv's strategy) ------------------------------------------
// Each evaluates its function to a relative error below FAST_EPS (about 2^-69 by construction: the
// leading series terms in double-double, the tail in double) and returns the rounded result only if
// both ends of [v - FAST_EPS |v|, v + FAST_EPS |v|] round to the same double; otherwise ok is false
// and the caller falls back to the full evaluation.  About one call in several thousand falls back.
constexpr double FAST_EPS = 0x1p-64;

// Double-double sum without the cancellation guard: its error is ~2^-104 of |a| + |b|, which is fine
// wherever the two operands do not nearly cancel (every use below).
CCD_FHD DD dd_add_fast(DD a, DD b)
{
    DD s = two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return quick_two_sum(s.hi, s.lo);
}
CCD_FHD double settle(DD v, bool &ok)
{
    double e = FAST_EPS * dabs(v.hi);
    double a = v.hi + (v.lo - e);
    double b = v.hi + (v.lo + e);
    ok = a == b;
    return a;
}

// log x = e ln2 + 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716; x positive, normal and finite.
CCD_FHD DD log_fast_dd(double x)
{
    long long bits = double_to_bits(x);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = bits_to_double((bits & 0x000fffffffffffffll) | 0x3ff0000000000000ll);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    double num = m - 1.0;            // exact (Sterbenz)
    DD den = two_sum(m, 1.0);        // m + 1 exactly
    double sh = num / den.hi;
    double rem = dfma(-sh, den.hi, num) - sh * den.lo;
    DD s = quick_two_sum(sh, rem / den.hi);
    DD s2 = two_prod(s.hi, s.hi);
    s2.lo += 2.0 * s.hi * s.lo;
    s2 = quick_two_sum(s2.hi, s2.lo);
    // sum_{k>=0} s^{2k}/(2k+1): k = 14..3 in double (the tail is below 2^-18), k = 2..0 in double-double.
    double t = INV_ODD[14][0];
#pragma unroll
    for (int k = 13; k >= 3; k--) t = dfma(t, s2.hi, INV_ODD[k][0]);
    DD acc = dd_add_fast(DD{INV_ODD[2][0], INV_ODD[2][1]}, DD{t * s2.hi, 0.0});
    acc = dd_add_fast(dd_mul(acc, s2), DD{INV_ODD[1][0], INV_ODD[1][1]});
    acc = dd_add_fast(dd_mul(acc, s2), DD{1.0, 0.0});
    DD lm = dd_mul(DD{2.0 * s.hi, 2.0 * s.lo}, acc);
    DD le = two_prod((double)e, LN2_HI);
    le.lo += (double)e * LN2_LO;
    return dd_add_fast(le, lm);
}

// exp y = 2^k exp(r), r = y - k ln2, |r| <= 0.35; y within [-708, 709.78].  Returns exp(r).
CCD_FHD DD exp_fast_dd(double y, int &k)
{
    double kd = y / LN2_HI;
    kd = kd < 0 ? -(double)(long long)(-kd + 0.5) : (double)(long long)(kd + 0.5);
    k = (int)kd;
    DD p = two_prod(kd, LN2_HI);
    DD r = two_sum(y - p.hi, -p.lo);   // y - p.hi is exact (Sterbenz)
    r.lo -= kd * LN2_LO;
    r = quick_two_sum(r.hi, r.lo);
    // sum_j r^j/j!: j = 17..6 in double (the tail is below 2^-18), j = 5..0 in double-double, at r.hi;
    // then exp(r) = exp(r.hi) (1 + r.lo) to the precision needed.
    double t = INV_FACT[17][0];
#pragma unroll
    for (int j = 16; j >= 6; j--) t = dfma(t, r.hi, INV_FACT[j][0]);
    DD acc = dd_add_fast(DD{INV_FACT[5][0], INV_FACT[5][1]}, DD{t * r.hi, 0.0});
#pragma unroll
    for (int j = 4; j >= 0; j--) acc = dd_add_fast(dd_mul_d(acc, r.hi), DD{INV_FACT[j][0], INV_FACT[j][1]});
    return dd_add_fast(acc, DD{acc.hi * r.lo, 0.0});
}

CCD_HD double log_cr(double x)
{
    if (x >= 0x1p-1000 && x < bits_to_double(0x7ff0000000000000ll)) {
        bool ok;
        double v = settle(log_fast_dd(x), ok);
        if (ok) return v;
    }
    return log_cr_full(x);
}

CCD_HD double exp_cr(double y)
{
    if (y >= -708.0 && y <= 709.78) {
        bool ok;
        int k;
        double v = settle(exp_fast_dd(y, k), ok);
        if (ok) return bits_ldexp(v, k);
    }
    return exp_cr_full(y);
}

// Test-only host build (CCD_HOST_LIBM): the solver calls the C library's log/exp instead.  The tests use
// it to show that a reference-mode result which differs from the reference's own solver differs only
// through the C library's rounding of log/exp (with them, the restatement is bit-identical).
#if defined(CCD_HOST_LIBM) && !defined(__HIP_DEVICE_COMPILE__)
CCD_HD double solver_log(double x) { return std::log(x); }
CCD_HD double solver_exp(double x) { return std::exp(x); }
#else
CCD_HD double solver_log(double x) { return log_cr(x); }
CCD_HD double solver_exp(double x) { return exp_cr(x); }
#endif

// ---- the three-stage algorithm (rpoly.h:49-759) -------------------------------------------------
constexpr double ETA = 2.22e-16;   // rpoly.h:57-60: machine constants as the reference sets them
constexpr double ARE = ETA, MRE = ETA;
constexpr double INFIN = 3.4e38, SMALNO = 1.2e-38;
constexpr double COSR = -0x1.1db902ec147ep-4;   // cos(94.0 * 0.017453293), rpoly.h:67-70
constexpr double SINR = 0x1.fec0b6fc05f2p-1;    // sin(94.0 * 0.017453293)

// The solver works on the live degree NN as a template parameter: between deflations the degree is
// fixed, so every routine below is instantiated per degree (3..6) and every array index is a
// compile-time constant -- on the GPU the arrays stay in registers with no index arithmetic.  Only the
// outer loop, which deflates, switches between the instantiations.  The arithmetic and its order are
// the reference's, unchanged.
struct Solver {
    double p[7] = {}, qp[7] = {}, k[7] = {}, qk[7] = {}, svk[7] = {};
    double sr = 0, si = 0, u = 0, v = 0, a = 0, b = 0, c = 0, d = 0, a1 = 0, a3 = 0, a7 = 0, e = 0, f = 0, g = 0,
           h = 0, szr = 0, szi = 0, lzr = 0, lzi = 0;
    int n = 0;

    // rpoly.h:665-680: divide pp by the quadratic 1,u,v; quotient to q, remainder to (a, b).
    template <int NN>
    CCD_FHD static void quadsd(double uu, double vv, const double (&pp)[7], double (&q)[7], double &aa, double &bb)
    {
        bb = pp[0];
        q[0] = bb;
        aa = pp[1] - bb * uu;
        q[1] = aa;
#pragma unroll
        for (int i = 2; i <= NN; i++) {
            double cc = pp[i] - aa * uu - bb * vv;
            q[i] = cc;
            bb = aa;
            aa = cc;
        }
    }

    // rpoly.h:687-733: zeros of a z^2 + b1 z + c.
    CCD_FHD static void quad(double a_, double b1, double c_, double &sr_, double &si_, double &lr, double &li)
    {
        double b_, d_, e_;
        if (a_ == 0.0) {
            sr_ = b1 != 0.0 ? -c_ / b1 : 0.0;
            lr = 0.0;
            si_ = 0.0;
            li = 0.0;
            return;
        }
        if (c_ == 0.0) {
            sr_ = 0.0;
            lr = -b1 / a_;
            si_ = 0.0;
            li = 0.0;
            return;
        }
        b_ = b1 / 2.0;
        if (dabs(b_) < dabs(c_)) {
            e_ = c_ < 0.0 ? -a_ : a_;
            e_ = b_ * (b_ / dabs(c_)) - e_;
            d_ = dsqrt(dabs(e_)) * dsqrt(dabs(c_));
        } else {
            e_ = 1.0 - (a_ / b_) * (c_ / b_);
            d_ = dsqrt(dabs(e_)) * dabs(b_);
        }
        if (e_ < 0.0) {
            sr_ = -b_ / a_;
            lr = sr_;
            si_ = dabs(d_ / a_);
            li = -si_;
        } else {
            if (b_ >= 0.0) d_ = -d_;
            lr = (-b_ + d_) / a_;
            sr_ = 0.0;
            if (lr != 0.0) sr_ = (c_ / lr) / a_;
            si_ = 0.0;
            li = 0.0;
        }
    }

    // rpoly.h:557-596: scalars for the next k polynomial; returns the normalisation type.
    template <int NN>
    CCD_FHD int calcsc()
    {
        quadsd<NN - 1>(u, v, k, qk, c, d);
        if (!(dabs(c) > dabs(k[NN - 1] * 100.0 * ETA)) && !(dabs(d) > dabs(k[NN - 2] * 100.0 * ETA))) return 3;
        if (dabs(d) < dabs(c)) {
            e = a / c;
            f = d / c;
            g = u * e;
            h = v * b;
            a3 = a * e + (h / c + g) * b;
            a1 = b - a * (d / c);
            a7 = a + g * d + h * f;
            return 1;
        }
        e = a / d;
        f = c / d;
        g = u * b;
        h = v * b;
        a3 = (a + g) * e + h * (b / d);
        a1 = b * f - a;
        a7 = (f + u) * a + h;
        return 2;
    }

    // rpoly.h:600-634
    template <int NN>
    CCD_FHD void nextk(int type)
    {
        if (type == 3) {
            k[0] = 0.0;
            k[1] = 0.0;
#pragma unroll
            for (int i = 2; i < NN; i++) k[i] = qk[i - 2];
            return;
        }
        double temp = type == 1 ? b : a;
        if (dabs(a1) <= dabs(temp) * ETA * 10.0) {
            k[0] = 0.0;
            k[1] = -a7 * qp[0];
#pragma unroll
            for (int i = 2; i < NN; i++) k[i] = a3 * qk[i - 2] - a7 * qp[i - 1];
            return;
        }
        a7 /= a1;
        a3 /= a1;
        k[0] = qp[0];
        k[1] = qp[1] - a7 * qp[0];
#pragma unroll
        for (int i = 2; i < NN; i++) k[i] = a3 * qk[i - 2] - a7 * qp[i - 1] + qp[i];
    }

    // rpoly.h:638-661
    template <int NN>
    CCD_FHD void newest(int type, double &uu, double &vv)
    {
        if (type == 3) {
            uu = 0.0;
            vv = 0.0;
            return;
        }
        double a4, a5;
        if (type == 2) {
            a4 = (a + g) * f + h;
            a5 = (f + u) * c + v * d;
        } else {
            a4 = a + u * b + h * f;
            a5 = c + (u + v * f) * d;
        }
        double b1 = -k[NN - 1] / p[NN];
        double b2 = -(k[NN - 2] + b1 * p[NN - 1]) / p[NN];
        double c1 = v * b2 * a1;
        double c2 = b1 * a7;
        double c3 = b1 * b1 * a3;
        double c4 = c1 - c2 - c3;
        double temp = a5 + b1 * a4 - c4;
        if (temp == 0.0) {
            uu = 0.0;
            vv = 0.0;
            return;
        }
        uu = u - (u * (c3 + c2) + v * (b1 * a1 + b2 * a7)) / temp;
        vv = v * (1.0 + c4 / temp);
    }

    // rpoly.h:401-479: variable-shift iteration for a quadratic factor; returns the zeros found (0 or 2).
    template <int NN>
    CCD_FHD int quadit(double uu, double vv)
    {
        double ui, vi, mp, omp = 0, ee, relstp = 0, t, zm;
        bool tried = false;
        int j = 0;
        u = uu;
        v = vv;
        while (true) {
            quad(1.0, u, v, szr, szi, lzr, lzi);
            if (dabs(dabs(szr) - dabs(lzr)) > 0.01 * dabs(lzr)) return 0;
            quadsd<NN>(u, v, p, qp, a, b);
            mp = dabs(a - szr * b) + dabs(szi * b);
            zm = dsqrt(dabs(v));
            ee = 2.0 * dabs(qp[0]);
            t = -szr * b;
#pragma unroll
            for (int i = 1; i < NN; i++) ee = ee * zm + dabs(qp[i]);
            ee = ee * zm + dabs(a + t);
            ee *= (5.0 * MRE + 4.0 * ARE);
            ee = ee - (5.0 * MRE + 2.0 * ARE) * (dabs(a + t) + dabs(b) * zm) + 2.0 * ARE * dabs(t);
            if (mp <= 20.0 * ee) return 2;
            j++;
            if (j > 20) return 0;
            if (!(j < 2) && !(relstp > 0.01 || mp < omp || tried)) {
                // A cluster stalls convergence: five fixed-shift steps near it.
                if (relstp < ETA) relstp = ETA;
                relstp = dsqrt(relstp);
                u = u - u * relstp;
                v = v + v * relstp;
                quadsd<NN>(u, v, p, qp, a, b);
                for (int i = 0; i < 5; i++) {
                    int type = calcsc<NN>();
                    nextk<NN>(type);
                }
                tried = true;
                j = 0;
            }
            omp = mp;
            int type = calcsc<NN>();
            nextk<NN>(type);
            type = calcsc<NN>();
            newest<NN>(type, ui, vi);
            if (vi == 0.0) return 0;
            relstp = dabs((vi - v) / vi);
            u = ui;
            v = vi;
        }
    }

    // rpoly.h:485-550: variable-shift iteration for a real zero; returns the zeros found (0 or 1) and
    // sets iflag (with s updated) when a cluster near the real axis calls for the quadratic iteration.
    template <int NN>
    CCD_FHD int realit(double &sss, int &iflag)
    {
        double pv, kv, t = 0, s = sss, ms, mp, omp = 0, ee;
        iflag = 0;
        int j = 0;
        while (true) {
            pv = p[0];
            qp[0] = pv;
#pragma unroll
            for (int i = 1; i <= NN; i++) {
                pv = pv * s + p[i];
                qp[i] = pv;
            }
            mp = dabs(pv);
            ms = dabs(s);
            ee = (MRE / (ARE + MRE)) * dabs(qp[0]);
#pragma unroll
            for (int i = 1; i <= NN; i++) ee = ee * ms + dabs(qp[i]);
            if (mp <= 20.0 * ((ARE + MRE) * ee - MRE * mp)) {
                szr = s;
                szi = 0.0;
                return 1;
            }
            j++;
            if (j > 10) return 0;
            if (!(j < 2) && !(dabs(t) > 0.001 * dabs(s - t) || mp < omp)) {
                iflag = 1;
                sss = s;
                return 0;
            }
            omp = mp;
            kv = k[0];
            qk[0] = kv;
#pragma unroll
            for (int i = 1; i < NN; i++) {
                kv = kv * s + k[i];
                qk[i] = kv;
            }
            if (dabs(kv) <= dabs(k[NN - 1]) * 10.0 * ETA) {
                k[0] = 0.0;
#pragma unroll
                for (int i = 1; i < NN; i++) k[i] = qk[i - 1];
            } else {
                t = -pv / kv;
                k[0] = qp[0];
#pragma unroll
                for (int i = 1; i < NN; i++) k[i] = t * qk[i - 1] + qp[i];
            }
            kv = k[0];
#pragma unroll
            for (int i = 1; i < NN; i++) kv = kv * s + k[i];
            t = 0.0;
            if (dabs(kv) > dabs(k[NN - 1] * 10.0 * ETA)) t = -pv / kv;
            s += t;
        }
    }

    // rpoly.h:278-395: up to l2 fixed-shift steps, then the variable-shift iteration of the faster
    // converging sequence; returns the number of zeros found.
    template <int NN>
    CCD_FHD int fxshfr(int l2)
    {
        double svu = 0, svv = 0, ui, vi, s = 0;
        double betas = 0.25, betav = 0.25, oss = sr, ovv = v, ss, vv, ts, tv, ots = 0, otv = 0, tvv, tss;
        int iflag;
        quadsd<NN>(u, v, p, qp, a, b);
        int type = calcsc<NN>();
        for (int j = 0; j < l2; j++) {
            nextk<NN>(type);
            type = calcsc<NN>();
            newest<NN>(type, ui, vi);
            vv = vi;
            ss = 0.0;
            if (k[NN - 1] != 0.0) ss = -p[NN] / k[NN - 1];
            tv = 1.0;
            ts = 1.0;
            if (!(j == 0 || type == 3)) {
                if (vv != 0.0) tv = dabs((vv - ovv) / vv);
                if (ss != 0.0) ts = dabs((ss - oss) / ss);
                tvv = 1.0;
                if (tv < otv) tvv = tv * otv;
                tss = 1.0;
                if (ts < ots) tss = ts * ots;
                bool vpass = tvv < betav;
                bool spass = tss < betas;
                if (spass || vpass) {
                    svu = u;
                    svv = v;
#pragma unroll
                    for (int i = 0; i < NN; i++) svk[i] = k[i];
                    s = ss;
                    bool vtry = false, stry = false;
                    enum { QUAD, LIN, RESTORE } step = ((spass && !vpass) || tss < tvv) ? LIN : QUAD;
                    while (true) {
                        if (step == QUAD) {
                            int nz = quadit<NN>(ui, vi);
                            if (nz > 0) return nz;
                            vtry = true;
                            betav *= 0.25;
                            if (stry || !spass) {
                                step = RESTORE;
                            } else {
#pragma unroll
                                for (int i = 0; i < NN; i++) k[i] = svk[i];
                                step = LIN;
                            }
                        } else if (step == LIN) {
                            int nz = realit<NN>(s, iflag);
                            if (nz > 0) return nz;
                            stry = true;
                            betas *= 0.25;
                            if (iflag == 0) {
                                step = RESTORE;
                            } else {
                                ui = -(s + s);
                                vi = s * s;
                                step = QUAD;
                            }
                        } else {
                            u = svu;
                            v = svv;
#pragma unroll
                            for (int i = 0; i < NN; i++) k[i] = svk[i];
                            if (vpass && !vtry) {
                                step = QUAD;
                                continue;
                            }
                            quadsd<NN>(u, v, p, qp, a, b);
                            type = calcsc<NN>();
                            break;
                        }
                    }
                }
            }
            ovv = vv;
            oss = ss;
            otv = tv;
            ots = ts;
        }
        return 0;
    }

    // One pass of the outer loop at degree NN (rpoly.h:119-247): scaling, the root-modulus bound, the
    // unshifted k steps, then up to 20 rotated shifts.  On success the zero(s) found are stored at
    // zeror/zeroi[degree - NN ..], the polynomial is deflated (n = NN - nz) and true is returned.
    template <int NN>
    CCD_FHD bool stage(int degree, double &xx, double &yy, double (&zeror)[6], double (&zeroi)[6])
    {
        const double lo = SMALNO / ETA;
        double temp[7] = {}, pt[7] = {};
        // Scale by a power of two if there are very large or very small coefficients.
        double mx = 0.0, mn = INFIN;
#pragma unroll
        for (int i = 0; i <= NN; i++) {
            double x = dabs(p[i]);
            if (x > mx) mx = x;
            if (x != 0.0 && x < mn) mn = x;
        }
        double sc = lo / mn;
        bool scale = true;
        if (sc > 1.0 && INFIN / sc < mx)
            scale = false;
        else if (sc <= 1.0) {
            if (mx < 10.0)
                scale = false;
            else if (sc == 0.0)
                sc = SMALNO;
        }
        if (scale) {
            int l = (int)(solver_log(sc) / solver_log(2.0) + 0.5);
            double factor = bits_ldexp(1.0, l);
            if (factor != 1.0) {
#pragma unroll
                for (int i = 0; i <= NN; i++) p[i] = factor * p[i];
            }
        }
        // Lower bound on the moduli of the zeros.
#pragma unroll
        for (int i = 0; i <= NN; i++) pt[i] = dabs(p[i]);
        pt[NN] = -pt[NN];
        double x = solver_exp((solver_log(-pt[NN]) - solver_log(pt[0])) / (double)NN);
        if (pt[NN - 1] != 0.0) {
            double xm = -pt[NN] / pt[NN - 1];
            if (xm < x) x = xm;
        }
        // Chop (0, x) until the bound polynomial is <= 0 (the cap only guards non-finite input, which
        // would never terminate).
        for (int guard = 0; guard < 1200; guard++) {
            double xm = x * 0.1;
            double ff = pt[0];
#pragma unroll
            for (int i = 1; i <= NN; i++) ff = ff * xm + pt[i];
            if (ff <= 0.0) break;
            x = xm;
        }
        double dx = x;
        for (int guard = 0; guard < 1000 && dabs(dx / x) > 0.005; guard++) {
            double ff = pt[0];
            double df = ff;
#pragma unroll
            for (int i = 1; i < NN; i++) {
                ff = ff * x + pt[i];
                df = df * x + ff;
            }
            ff = ff * x + pt[NN];
            dx = ff / df;
            x -= dx;
        }
        double bnd = x;
        // The derivative as the first k polynomial, then 5 unshifted steps.
#pragma unroll
        for (int i = 1; i < NN; i++) k[i] = (double)(NN - i) * p[i] / (double)NN;
        k[0] = p[0];
        double aa = p[NN], bb = p[NN - 1];
        bool zerok = k[NN - 1] == 0;
        for (int jj = 0; jj < 5; jj++) {
            double cc = k[NN - 1];
            if (!zerok) {
                double t = -aa / cc;
                // j = NN-1 down to 1, each reading the not yet updated k[j-1]
#pragma unroll
                for (int j = NN - 1; j >= 1; j--) k[j] = t * k[j - 1] + p[j];
                k[0] = p[0];
                zerok = dabs(k[NN - 1]) <= dabs(bb) * ETA * 10.0;
            } else {
#pragma unroll
                for (int j = NN - 1; j >= 1; j--) k[j] = k[j - 1];
                k[0] = 0.0;
                zerok = k[NN - 1] == 0.0;
            }
        }
#pragma unroll
        for (int i = 0; i < NN; i++) temp[i] = k[i];
        // Up to 20 shifts, each rotated by 94 degrees from the last.
        for (int cnt = 0; cnt < 20; cnt++) {
            double xxx = COSR * xx - SINR * yy;
            yy = SINR * xx + COSR * yy;
            xx = xxx;
            sr = bnd * xx;
            si = bnd * yy;
            u = -2.0 * sr;
            v = bnd;
            int nz = fxshfr<NN>(20 * (cnt + 1));
            if (nz != 0) {
                int j = degree - NN;
                put(zeror, j, szr);
                put(zeroi, j, szi);
                n = NN - nz;
#pragma unroll
                for (int i = 0; i <= NN; i++)
                    if (i <= NN - nz) p[i] = qp[i];
                if (nz != 1) {
                    put(zeror, j + 1, lzr);
                    put(zeroi, j + 1, lzi);
                }
                return true;
            }
#pragma unroll
            for (int i = 0; i < NN; i++) k[i] = temp[i];
        }
        return false;   // no convergence with 20 shifts
    }

    // rpoly.h:49-261.  op[0..degree] descending, degree <= 6 (slots above degree are ignored).  Returns
    // the number of zeros found (degree, or fewer if 20 shifts fail), or -1 for a zero leading
    // coefficient or a polynomial that is a monomial.
    CCD_FHD int rpoly(const double (&op)[7], int degree, double (&zeror)[6], double (&zeroi)[6])
    {
        double xx = 0.70710678118654757;   // sqrt(0.5), correctly rounded
        double yy = -xx;
        n = degree;
        if (op[0] == 0.0) return -1;
        while (sel(op, n) == 0.0) {
            int j = degree - n;
            put(zeror, j, 0.0);
            put(zeroi, j, 0.0);
            n--;
        }
        if (n < 1) return -1;
#pragma unroll
        for (int i = 0; i < 7; i++) p[i] = i <= n ? op[i] : 0.0;
        while (true) {
            if (n == 1) {
                put(zeror, degree - 1, -p[1] / p[0]);
                put(zeroi, degree - 1, 0.0);
                n -= 1;
                break;
            }
            if (n == 2) {
                double r0, i0, r1, i1;
                quad(p[0], p[1], p[2], r0, i0, r1, i1);
                put(zeror, degree - 2, r0);
                put(zeroi, degree - 2, i0);
                put(zeror, degree - 1, r1);
                put(zeroi, degree - 1, i1);
                n -= 2;
                break;
            }
            bool found = false;
            if (n == 3) found = stage<3>(degree, xx, yy, zeror, zeroi);
            else if (n == 4) found = stage<4>(degree, xx, yy, zeror, zeroi);
            else if (n == 5) found = stage<5>(degree, xx, yy, zeror, zeroi);
            else if (n == 6) found = stage<6>(degree, xx, yy, zeror, zeroi);
            if (!found) break;
        }
        return degree - n;
    }
};

// The solver as one out-of-line function: its state stays in this function's registers, and the
// narrowphase kernels that call it keep their own register budget.
static __attribute__((noinline)) CCD_HD_NI int find_roots(const double (&op)[7], int degree, double (&zeror)[6],
                                                  double (&zeroi)[6])
{
    Solver js;
    return js.rpoly(op, degree, zeror, zeroi);
}

// std::sort of at most 6 doubles as libstdc++ does it for short ranges (insertion sort whose
// element smaller than the first goes straight to the front), with compile-time indices.
CCD_HD void sort_small(double (&x)[6], int n)
{
#pragma unroll
    for (int i = 1; i < 6; i++) {
        if (i < n) {
            double val = x[i];
            if (val < x[0]) {
#pragma unroll
                for (int j = i; j >= 1; j--) x[j] = x[j - 1];
                x[0] = val;
            } else {
                // unguarded linear insert: shift while val < x[j-1]; x[0] <= val stops it
                bool moving = true;
#pragma unroll
                for (int j = i; j >= 1; j--) {
                    if (moving) {
                        if (val < x[j - 1]) {
                            x[j] = x[j - 1];
                        } else {
                            x[j] = val;
                            moving = false;
                        }
                    }
                }
                if (moving) x[0] = val;
            }
        }
    }
}

}   // namespace jt
}   // namespace ccd
