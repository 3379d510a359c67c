// Separating-plane narrowphase (reference: src/SeparatingPlaneNarrowPhase.h:7-22,
// src/SeparatingPlaneNarrowPhase.cpp), computed on the GPU by ccd_separating_plane_narrowphase.  As in
// the reference, colliding stencils are inserted into vfs / ees, which are not cleared first.  Where
// the reference's recursion would never end (it overflows its stack there) this throws instead.
#pragma once

#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "Device.h"
#include "RetrospectiveDetection.h"

class SeparatingPlaneNarrowPhase : public NarrowPhase {
public:
    void findCollisions(const History &h, const std::set<std::pair<VertexFaceStencil, double>> &candidateVFS,
                        const std::set<std::pair<EdgeEdgeStencil, double>> &candidateEES,
                        std::set<VertexFaceStencil> &vfs, std::set<EdgeEdgeStencil> &ees) override
    {
        const int64_t nvf = (int64_t)candidateVFS.size(), nee = (int64_t)candidateEES.size();
        if (nvf == 0 && nee == 0) return;
        std::vector<int32_t> vk, ek;
        std::vector<double> veta, eeta;
        vk.reserve(4 * nvf);
        ek.reserve(4 * nee);
        for (const auto &c : candidateVFS) {
            vk.insert(vk.end(), {c.first.p, c.first.q0, c.first.q1, c.first.q2});
            veta.push_back(c.second);
        }
        for (const auto &c : candidateEES) {
            ek.insert(ek.end(), {c.first.p0, c.first.p1, c.first.q0, c.first.q1});
            eeta.push_back(c.second);
        }
        ccdamd::Device &dev = ccdamd::Device::get();
        std::lock_guard<std::mutex> lock(dev.mutex());
        ccd_ctx *ctx = dev.ctx();
        std::vector<int64_t> off;
        std::vector<double> ht, hp;
        h.csr(off, ht, hp);
        ccdamd::DeviceBuffer o(off), t(ht), p(hp);
        ccdamd::DeviceBuffer dvk(vk), dek(ek), dve(veta), dee(eeta);
        ccdamd::DeviceBuffer vst(nvf + 1), est(nee + 1);
        int64_t undecided = 0;
        ccdamd::check(ctx,
                      ccd_separating_plane_narrowphase(ctx, o.as<int64_t>(), t.as<double>(), p.as<double>(),
                                                       h.numVertices(), (int64_t)ht.size(), dvk.as<int32_t>(),
                                                       dve.as<double>(), nvf, dek.as<int32_t>(), dee.as<double>(), nee,
                                                       0.0, vst.as<uint8_t>(), est.as<uint8_t>(), &undecided),
                      "SeparatingPlaneNarrowPhase::findCollisions");
        if (undecided)
            throw std::runtime_error("SeparatingPlaneNarrowPhase::findCollisions: the recursion does not terminate for " +
                                     std::to_string(undecided) + " stencil(s)");
        std::vector<uint8_t> hv(nvf), he(nee);
        vst.download(hv.data(), nvf);
        est.download(he.data(), nee);
        int64_t i = 0;
        for (const auto &c : candidateVFS)
            if (hv[i++] == 1) vfs.insert(c.first);
        i = 0;
        for (const auto &c : candidateEES)
            if (he[i++] == 1) ees.insert(c.first);
    }
};
