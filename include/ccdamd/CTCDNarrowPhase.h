// Continuous-time narrowphase (reference: src/CTCDNarrowPhase.h:6-15, src/CTCDNarrowPhase.cpp), computed
// on the GPU by ccd_ctcd_narrowphase / ccd_ctcd_narrowphase_history.  As in the reference, colliding
// stencils are inserted into vfs / ees, which are not cleared first.
#pragma once

#include <set>
#include <stdexcept>
#include <vector>

#include "Device.h"
#include "RetrospectiveDetection.h"

class CTCDNarrowPhase : public NarrowPhase {
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
        const int64_t V = h.numVertices();
        ccdamd::DeviceBuffer dvk(vk), dek(ek), dve(veta), dee(eeta);
        ccdamd::DeviceBuffer vhit(nvf + 1), ehit(nee + 1);
        if (h.twoEntry()) {
            std::vector<double> qs, qe;
            h.endpoints(qs, qe);
            ccdamd::DeviceBuffer a(qs), b(qe);
            ccdamd::check(ctx,
                          ccd_ctcd_narrowphase(ctx, a.as<double>(), b.as<double>(), V, dvk.as<int32_t>(), dve.as<double>(),
                                               nvf, dek.as<int32_t>(), dee.as<double>(), nee, 0.0, vhit.as<uint8_t>(),
                                               nullptr, nullptr, ehit.as<uint8_t>(), nullptr, nullptr),
                          "CTCDNarrowPhase::findCollisions");
        } else {
            std::vector<int64_t> off;
            std::vector<double> ht, hp;
            h.csr(off, ht, hp);
            ccdamd::DeviceBuffer o(off), t(ht), p(hp);
            ccdamd::check(ctx,
                          ccd_ctcd_narrowphase_history(ctx, o.as<int64_t>(), t.as<double>(), p.as<double>(), V,
                                                       (int64_t)ht.size(), dvk.as<int32_t>(), dve.as<double>(), nvf,
                                                       dek.as<int32_t>(), dee.as<double>(), nee, 0.0,
                                                       vhit.as<uint8_t>(), nullptr, nullptr, nullptr,
                                                       ehit.as<uint8_t>(), nullptr, nullptr, nullptr),
                          "CTCDNarrowPhase::findCollisions");
        }
        std::vector<uint8_t> hv(nvf), he(nee);
        vhit.download(hv.data(), nvf);
        ehit.download(he.data(), nee);
        int64_t i = 0;
        for (const auto &c : candidateVFS)
            if (hv[i++]) vfs.insert(c.first);
        i = 0;
        for (const auto &c : candidateEES)
            if (he[i++]) ees.insert(c.first);
    }
};
