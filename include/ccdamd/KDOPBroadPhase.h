// k-DOP broadphase (reference: src/KDOPBroadPhase.h:49-61, src/KDOPBroadPhase.cpp:34-185), computed on the
// GPU by ccd_kdop_candidates / ccd_kdop_candidates_history.  Like the reference, it clears the output
// sets and fills them with the unique candidate stencils of every overlapping, non-neighbouring face
// pair (13-axis swept k-DOPs grown by outerEta), skipping stencils whose 4 vertices are all fixed.
#pragma once

#include <set>
#include <vector>

#include "Device.h"
#include "Mesh.h"
#include "RetrospectiveDetection.h"

class KDOPBroadPhase : public BroadPhase {
public:
    KDOPBroadPhase() {}

    void findCollisionCandidates(const History &h, const Mesh &m, double outerEta, std::set<VertexFaceStencil> &vfs,
                                 std::set<EdgeEdgeStencil> &ees, const std::set<int> &fixedVerts) override
    {
        vfs.clear();
        ees.clear();
        ccdamd::Device &dev = ccdamd::Device::get();
        std::lock_guard<std::mutex> lock(dev.mutex());
        ccd_ctx *ctx = dev.ctx();
        const int64_t F = m.faces.cols();
        const int64_t V = h.numVertices();
        ccdamd::DeviceBuffer faces(m.faces.data(), F * 3 * (int64_t)sizeof(int));
        std::vector<uint8_t> fx;
        ccdamd::DeviceBuffer fixed;
        if (!fixedVerts.empty()) {
            fx.assign(V, 0);
            for (int v : fixedVerts)
                if (v >= 0 && v < V) fx[v] = 1;
            fixed.upload(fx.data(), V);
        }
        std::vector<double> qs, qe, ht, hp;
        std::vector<int64_t> off;
        ccdamd::DeviceBuffer a, b, c;
        bool two = h.twoEntry();
        if (two) {
            h.endpoints(qs, qe);
            a.upload(qs.data(), (int64_t)qs.size() * 8);
            b.upload(qe.data(), (int64_t)qe.size() * 8);
        } else {
            h.csr(off, ht, hp);
            a.upload(off.data(), (int64_t)off.size() * 8);
            c.upload(hp.data(), (int64_t)hp.size() * 8);
        }
        int64_t cap_v = 0, cap_e = 0, nv = 0, ne = 0;
        ccdamd::DeviceBuffer vout, eout;
        for (int attempt = 0; attempt < 2; attempt++) {
            vout.resize(16 * (cap_v + 1));
            eout.resize(16 * (cap_e + 1));
            int rc = two ? ccd_kdop_candidates(ctx, faces.as<int32_t>(), F, a.as<double>(), b.as<double>(), V, outerEta,
                                               fixedVerts.empty() ? nullptr : fixed.as<uint8_t>(), vout.as<int32_t>(),
                                               cap_v, &nv, eout.as<int32_t>(), cap_e, &ne, CCD_SORTED)
                         : ccd_kdop_candidates_history(ctx, faces.as<int32_t>(), F, a.as<int64_t>(), c.as<double>(), V,
                                                       (int64_t)ht.size(), outerEta,
                                                       fixedVerts.empty() ? nullptr : fixed.as<uint8_t>(),
                                                       vout.as<int32_t>(), cap_v, &nv, eout.as<int32_t>(), cap_e, &ne,
                                                       CCD_SORTED);
            if (rc == CCD_E_CAPACITY && attempt == 0) {
                cap_v = nv;
                cap_e = ne;
                continue;
            }
            ccdamd::check(ctx, rc, "KDOPBroadPhase::findCollisionCandidates");
            break;
        }
        std::vector<int32_t> hv(4 * nv), he(4 * ne);
        vout.download(hv.data(), 16 * nv);
        eout.download(he.data(), 16 * ne);
        // Keys arrive in set order: hinted insertion at the end is constant time each.
        for (int64_t i = 0; i < nv; i++)
            vfs.insert(vfs.end(), VertexFaceStencil(hv[4 * i], hv[4 * i + 1], hv[4 * i + 2], hv[4 * i + 3]));
        for (int64_t i = 0; i < ne; i++)
            ees.insert(ees.end(), EdgeEdgeStencil(he[4 * i], he[4 * i + 1], he[4 * i + 2], he[4 * i + 3]));
    }
};
