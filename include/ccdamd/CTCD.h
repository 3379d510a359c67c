// Conservative continuous-time primitive tests (reference: include/CTCD.h, src/CTCD.cpp:259-692),
// computed on the GPU by ccd_ctcd_primitives.  Same signatures: true on a collision closer than eta,
// with the earliest time in [0,1] written to t (t is untouched otherwise).  Each call is one device
// round trip; the *Batch functions test many primitives in one launch.
#pragma once

#include <algorithm>
#include <vector>

#include "Device.h"
#include "EigenTypes.h"

struct TimeInterval {
    // CTCD.h:9-14: order the ends, clamp to [0,1].
    TimeInterval(double tl, double tu) : l(tl), u(tu)
    {
        if (l > u) std::swap(l, u);
        l = std::max(l, 0.0);
        u = std::min(u, 1.0);
    }
    TimeInterval() : l(0), u(0) {}
    // CTCD.cpp:10-36
    static bool overlap(const TimeInterval &a, const TimeInterval &b) { return !(a.l > b.u || b.l > a.u); }
    static bool overlap(const std::vector<TimeInterval> &iv)
    {
        for (size_t i = 0; i < iv.size(); i++)
            for (size_t j = i + 1; j < iv.size(); j++)
                if (!overlap(iv[i], iv[j])) return false;
        return true;
    }
    static TimeInterval intersect(const std::vector<TimeInterval> &iv)
    {
        TimeInterval r(0.0, 1.0);
        for (const auto &x : iv) {
            r.l = std::max(x.l, r.l);
            r.u = std::min(x.u, r.u);
        }
        return r;
    }
    double l, u;
};

class CTCD {
public:
    static bool edgeEdgeCTCD(const Eigen::Vector3d &q0start, const Eigen::Vector3d &p0start,
                             const Eigen::Vector3d &q1start, const Eigen::Vector3d &p1start,
                             const Eigen::Vector3d &q0end, const Eigen::Vector3d &p0end, const Eigen::Vector3d &q1end,
                             const Eigen::Vector3d &p1end, double eta, double &t)
    {
        const Eigen::Vector3d *p[8] = {&q0start, &p0start, &q1start, &p1start, &q0end, &p0end, &q1end, &p1end};
        return one(CCD_PRIM_EDGE_EDGE, p, 8, eta, t);
    }
    static bool vertexFaceCTCD(const Eigen::Vector3d &q0start, const Eigen::Vector3d &q1start,
                               const Eigen::Vector3d &q2start, const Eigen::Vector3d &q3start,
                               const Eigen::Vector3d &q0end, const Eigen::Vector3d &q1end,
                               const Eigen::Vector3d &q2end, const Eigen::Vector3d &q3end, double eta, double &t)
    {
        const Eigen::Vector3d *p[8] = {&q0start, &q1start, &q2start, &q3start, &q0end, &q1end, &q2end, &q3end};
        return one(CCD_PRIM_VERTEX_FACE, p, 8, eta, t);
    }
    static bool vertexEdgeCTCD(const Eigen::Vector3d &q0start, const Eigen::Vector3d &q1start,
                               const Eigen::Vector3d &q2start, const Eigen::Vector3d &q0end,
                               const Eigen::Vector3d &q1end, const Eigen::Vector3d &q2end, double eta, double &t)
    {
        const Eigen::Vector3d *p[6] = {&q0start, &q1start, &q2start, &q0end, &q1end, &q2end};
        return one(CCD_PRIM_VERTEX_EDGE, p, 6, eta, t);
    }
    static bool vertexVertexCTCD(const Eigen::Vector3d &q1start, const Eigen::Vector3d &q2start,
                                 const Eigen::Vector3d &q1end, const Eigen::Vector3d &q2end, double eta, double &t)
    {
        const Eigen::Vector3d *p[4] = {&q1start, &q2start, &q1end, &q2end};
        return one(CCD_PRIM_VERTEX_VERTEX, p, 4, eta, t);
    }

    // n primitives of one kind in one launch: q holds each item's start positions then end positions
    // (24/24/18/12 doubles per item, see ccd_ctcd_primitives); eta per item or one value.
    static void batch(int kind, const std::vector<double> &q, const std::vector<double> &eta, std::vector<uint8_t> &hit,
                      std::vector<double> &t)
    {
        static const int width[4] = {24, 24, 18, 12};
        const int64_t n = (int64_t)q.size() / width[kind];
        hit.assign(n, 0);
        t.assign(n, 0.0);
        if (n == 0) return;
        ccdamd::Device &dev = ccdamd::Device::get();
        std::lock_guard<std::mutex> lock(dev.mutex());
        ccd_ctx *ctx = dev.ctx();
        ccdamd::DeviceBuffer dq(q), dh(n), dt(8 * n);
        ccdamd::DeviceBuffer de;
        bool per_item = (int64_t)eta.size() == n && n > 1;
        if (per_item) de.upload(eta.data(), 8 * n);
        ccdamd::check(ctx,
                      ccd_ctcd_primitives(ctx, kind, dq.as<double>(), per_item ? de.as<double>() : nullptr,
                                          eta.empty() ? 0.0 : eta[0], n, dh.as<uint8_t>(), dt.as<double>()),
                      "CTCD::batch");
        dh.download(hit.data(), n);
        dt.download(t.data(), 8 * n);
    }

private:
    static bool one(int kind, const Eigen::Vector3d *const *p, int np, double eta, double &t)
    {
        std::vector<double> q(3 * np);
        for (int i = 0; i < np; i++)
            for (int k = 0; k < 3; k++) q[3 * i + k] = (*p[i])[k];
        std::vector<uint8_t> hit;
        std::vector<double> tt;
        batch(kind, q, std::vector<double>(1, eta), hit, tt);
        if (hit[0]) t = tt[0];
        return hit[0] != 0;
    }
};
