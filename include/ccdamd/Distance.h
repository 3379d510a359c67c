// Proximity statics (reference: include/Distance.h:9-188, src/Distance.cpp:12-66), computed on the GPU by
// ccd_distance_primitives and ccd_mesh_self_distance.  Same signatures and outputs.  Each static is one
// device round trip; Distance::batch tests many items of one kind in one launch.  meshSelfDistance
// prints the candidate counts the reference prints (Distance.cpp:47).
#pragma once

#include <cstdint>
#include <iostream>
#include <limits>
#include <set>
#include <vector>

#include "Device.h"
#include "EigenTypes.h"

class Distance {
public:
    static bool vertexPlaneDistanceLessThan(const Eigen::Vector3d &p, const Eigen::Vector3d &q0,
                                            const Eigen::Vector3d &q1, const Eigen::Vector3d &q2, double eta)
    {
        return one(CCD_DIST_VERTEX_PLANE_LT, p, q0, q1, q2, eta, nullptr, nullptr);
    }
    static bool lineLineDistanceLessThan(const Eigen::Vector3d &p0, const Eigen::Vector3d &p1,
                                         const Eigen::Vector3d &q0, const Eigen::Vector3d &q1, double eta)
    {
        return one(CCD_DIST_LINE_LINE_LT, p0, p1, q0, q1, eta, nullptr, nullptr);
    }
    static Eigen::Vector3d vertexFaceDistance(const Eigen::Vector3d &p, const Eigen::Vector3d &q0,
                                              const Eigen::Vector3d &q1, const Eigen::Vector3d &q2, double &q0bary,
                                              double &q1bary, double &q2bary)
    {
        double vec[3], bary[4];
        one(CCD_DIST_VERTEX_FACE, p, q0, q1, q2, 0.0, vec, bary);
        q0bary = bary[0];
        q1bary = bary[1];
        q2bary = bary[2];
        return Eigen::Vector3d(vec[0], vec[1], vec[2]);
    }
    static Eigen::Vector3d edgeEdgeDistance(const Eigen::Vector3d &p0, const Eigen::Vector3d &p1,
                                            const Eigen::Vector3d &q0, const Eigen::Vector3d &q1, double &p0bary,
                                            double &p1bary, double &q0bary, double &q1bary)
    {
        double vec[3], bary[4];
        one(CCD_DIST_EDGE_EDGE, p0, p1, q0, q1, 0.0, vec, bary);
        p0bary = bary[0];
        p1bary = bary[1];
        q0bary = bary[2];
        q1bary = bary[3];
        return Eigen::Vector3d(vec[0], vec[1], vec[2]);
    }

    // Distance.cpp:12-66: the shortest vertex-face / edge-edge distance between primitives sharing no
    // vertex, ignoring those whose vertices are all in fixedVerts.
    static double meshSelfDistance(const Eigen::VectorXd &verts, const Eigen::Matrix3Xi &faces,
                                   const std::set<int> &fixedVerts)
    {
        const int64_t V = (int64_t)verts.size() / 3, F = (int64_t)faces.cols();
        std::vector<double> q(3 * V);
        for (int64_t i = 0; i < 3 * V; i++) q[i] = verts[i];
        std::vector<int32_t> f(3 * F);
        for (int64_t c = 0; c < F; c++)
            for (int r = 0; r < 3; r++) f[3 * c + r] = faces.coeff(r, c);
        ccdamd::Device &dev = ccdamd::Device::get();
        std::lock_guard<std::mutex> lock(dev.mutex());
        ccd_ctx *ctx = dev.ctx();
        ccdamd::DeviceBuffer dq(q), df(f), dfx;
        if (!fixedVerts.empty()) {
            std::vector<uint8_t> fx(V, 0);
            for (int v : fixedVerts)
                if (v >= 0 && v < V) fx[v] = 1;
            dfx.upload(fx.data(), V);
        }
        double d = std::numeric_limits<double>::infinity();
        int64_t nvf = 0, nee = 0;
        ccdamd::check(ctx,
                      ccd_mesh_self_distance(ctx, dq.as<double>(), V, df.as<int32_t>(), F,
                                             fixedVerts.empty() ? nullptr : dfx.as<uint8_t>(), &d, &nvf, &nee),
                      "Distance::meshSelfDistance");
        std::cout << "Checking " << nvf << " vertex-face and " << nee << " edge-edge stencils" << std::endl;
        return d;
    }

    // n items of one kind (CCD_DIST_*) in one launch: q holds 12 doubles per item (4 points in the
    // argument order of the static); flag (LessThan kinds) or vec (3 per item) and bary (4 per item).
    static void batch(int kind, const std::vector<double> &q, double eta, std::vector<uint8_t> &flag,
                      std::vector<double> &vec, std::vector<double> &bary)
    {
        const int64_t n = (int64_t)q.size() / 12;
        flag.assign(n, 0);
        vec.assign(3 * n, 0.0);
        bary.assign(4 * n, 0.0);
        if (n == 0) return;
        ccdamd::Device &dev = ccdamd::Device::get();
        std::lock_guard<std::mutex> lock(dev.mutex());
        ccd_ctx *ctx = dev.ctx();
        ccdamd::DeviceBuffer dq(q), dfl(n), dv(24 * n), db(32 * n);
        ccdamd::check(ctx,
                      ccd_distance_primitives(ctx, kind, dq.as<double>(), eta, n, dfl.as<uint8_t>(), dv.as<double>(),
                                              db.as<double>()),
                      "Distance::batch");
        dfl.download(flag.data(), n);
        dv.download(vec.data(), 24 * n);
        db.download(bary.data(), 32 * n);
    }

private:
    static bool one(int kind, const Eigen::Vector3d &a, const Eigen::Vector3d &b, const Eigen::Vector3d &c,
                    const Eigen::Vector3d &d, double eta, double *vec, double *bary)
    {
        std::vector<double> q(12);
        const Eigen::Vector3d *p[4] = {&a, &b, &c, &d};
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) q[3 * i + k] = (*p[i])[k];
        std::vector<uint8_t> fl;
        std::vector<double> v, bb;
        batch(kind, q, eta, fl, v, bb);
        if (vec)
            for (int k = 0; k < 3; k++) vec[k] = v[k];
        if (bary)
            for (int k = 0; k < 4; k++) bary[k] = bb[k];
        return fl[0] != 0;
    }
    static double clamp(double u) { return std::min(1.0, std::max(u, 0.0)); }   // Distance.h:183-186
};
