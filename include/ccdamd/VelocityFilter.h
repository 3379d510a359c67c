// Velocity filter (reference: include/VelocityFilter.h:7-45, src/VelocityFilter.cpp:16-87), computed by
// ccd_velocity_filter: the layered penalty integration of src/ActiveLayers.cpp, src/PenaltyGroup.cpp and
// src/PenaltyPotential.cpp on the host, meshSelfDistance, the k-DOP broadphase over the multi-entry
// history and the separating-plane narrowphase on the GPU.  Same signature, defaults, return codes and
// progress lines on stdout; qend is overwritten only on success.  Where the separating-plane recursion
// never terminates for a stencil (the reference overflows its stack there) this throws instead.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "Device.h"
#include "EigenTypes.h"

class VelocityFilter {
public:
    static constexpr int DANGLING_VERTICES = -1;
    static constexpr int STARTING_STATE_INTERSECTS = -2;
    static constexpr int MAXROLLBACKS_EXCEEDED = -3;

    // masses: inverse masses per DOF (3 per vertex; 0 = infinite mass), as in the reference.
    static int velocityFilter(const Eigen::VectorXd &qstart, Eigen::VectorXd &qend, const Eigen::Matrix3Xi &faces,
                              const Eigen::VectorXd &masses, double outerEta, double innerEta,
                              double baseStiffness = .0001, double baseSubstepSize = 0.0001, double CoR = 0.1,
                              int maxRollbacks = 1000)
    {
        ccd_vf_stats st;
        return velocityFilter(qstart, qend, faces, masses, outerEta, innerEta, baseStiffness, baseSubstepSize, CoR,
                              maxRollbacks, st, true);
    }

    // The same, with the call's counters and the progress lines optional.
    static int velocityFilter(const Eigen::VectorXd &qstart, Eigen::VectorXd &qend, const Eigen::Matrix3Xi &faces,
                              const Eigen::VectorXd &masses, double outerEta, double innerEta, double baseStiffness,
                              double baseSubstepSize, double CoR, int maxRollbacks, ccd_vf_stats &stats,
                              bool verbose)
    {
        const int64_t n = (int64_t)qstart.size(), F = (int64_t)faces.cols();
        if ((int64_t)qend.size() != n || (int64_t)masses.size() != n || n % 3 != 0)
            throw std::invalid_argument("VelocityFilter::velocityFilter: qstart, qend and masses need 3 entries per vertex");
        std::vector<double> qs(n), qe(n), im(n);
        for (int64_t i = 0; i < n; i++) {
            qs[i] = qstart[i];
            qe[i] = qend[i];
            im[i] = masses[i];
        }
        std::vector<int32_t> f(3 * F);
        for (int64_t c = 0; c < F; c++)
            for (int r = 0; r < 3; r++) f[3 * c + r] = faces.coeff(r, c);
        ccdamd::Device &dev = ccdamd::Device::get();
        std::lock_guard<std::mutex> lock(dev.mutex());
        int result = 0;
        ccdamd::check(dev.ctx(),
                      ccd_velocity_filter(dev.ctx(), qs.data(), qe.data(), n / 3, f.data(), F, im.data(), outerEta,
                                          innerEta, baseStiffness, baseSubstepSize, CoR, maxRollbacks, verbose ? 1 : 0,
                                          &result, &stats),
                      "VelocityFilter::velocityFilter");
        if (result >= 0)
            for (int64_t i = 0; i < n; i++) qend[i] = qe[i];
        return result;
    }
};
