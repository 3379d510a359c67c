// Test driver for the C++ interface (include/ccdamd/), written the way the reference's AlecTest driver
// uses the library (example/AlecTest.cpp:82-115): History + finishHistory, KDOPBroadPhase, per-stencil
// eta, CTCDNarrowPhase and SeparatingPlaneNarrowPhase, plus the CTCD statics on testCTCD's four cases.
// Usage: cpp_interface_driver <faces.i32> <qstart.f64> <qend.f64> <eta> <out.txt>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "CTCD.h"
#include "CTCDNarrowPhase.h"
#include "History.h"
#include "KDOPBroadPhase.h"
#include "Mesh.h"
#include "SeparatingPlaneNarrowPhase.h"

template <class T>
static std::vector<T> read_all(const char *path)
{
    std::ifstream f(path, std::ios::binary);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<T> v(b.size() / sizeof(T));
    std::copy(b.begin(), b.begin() + v.size() * sizeof(T), (char *)v.data());
    return v;
}

int main(int argc, char **argv)
{
    if (argc != 6) {
        std::cerr << "usage: cpp_interface_driver faces qstart qend eta out" << std::endl;
        return 2;
    }
    std::vector<int> fi = read_all<int>(argv[1]);
    std::vector<double> qs = read_all<double>(argv[2]), qe = read_all<double>(argv[3]);
    double eta = atof(argv[4]);
    Eigen::VectorXd initialv((std::ptrdiff_t)qs.size()), finalv((std::ptrdiff_t)qe.size());
    for (size_t i = 0; i < qs.size(); i++) initialv[i] = qs[i];
    for (size_t i = 0; i < qe.size(); i++) finalv[i] = qe[i];
    Mesh m;
    m.faces.resize(3, (std::ptrdiff_t)(fi.size() / 3));
    for (size_t f = 0; f < fi.size() / 3; f++)
        for (int j = 0; j < 3; j++) m.faces.coeffRef(j, f) = fi[3 * f + j];
    m.vertices = initialv;
    History h(initialv);
    h.finishHistory(finalv);

    std::set<VertexFaceStencil> can_vfs;
    std::set<EdgeEdgeStencil> can_ees;
    const std::set<int> fixedVerts;
    KDOPBroadPhase().findCollisionCandidates(h, m, eta, can_vfs, can_ees, fixedVerts);
    std::set<std::pair<VertexFaceStencil, double>> vfs_eta;
    std::set<std::pair<EdgeEdgeStencil, double>> ees_eta;
    for (const auto &s : can_vfs) vfs_eta.insert(std::make_pair(s, eta));
    for (const auto &s : can_ees) ees_eta.insert(std::make_pair(s, eta));
    std::set<VertexFaceStencil> vfs;
    std::set<EdgeEdgeStencil> ees;
    CTCDNarrowPhase().findCollisions(h, vfs_eta, ees_eta, vfs, ees);

    std::ofstream out(argv[5]);
    out << "candidates " << can_vfs.size() << " " << can_ees.size() << "\n";
    out << "collisions " << vfs.size() << " " << ees.size() << "\n";
    for (const auto &s : vfs) out << "vf " << s.p << " " << s.q0 << " " << s.q1 << " " << s.q2 << "\n";
    for (const auto &s : ees) out << "ee " << s.p0 << " " << s.p1 << " " << s.q0 << " " << s.q1 << "\n";
    std::set<VertexFaceStencil> spvfs;
    std::set<EdgeEdgeStencil> spees;
    SeparatingPlaneNarrowPhase().findCollisions(h, vfs_eta, ees_eta, spvfs, spees);
    out << "sp_collisions " << spvfs.size() << " " << spees.size() << "\n";
    for (const auto &s : spvfs) out << "spvf " << s.p << " " << s.q0 << " " << s.q1 << " " << s.q2 << "\n";
    for (const auto &s : spees) out << "spee " << s.p0 << " " << s.p1 << " " << s.q0 << " " << s.q1 << "\n";

    // example/testCTCD.cpp's four cases (eta = 1e-6)
    double t = -1;
    bool r;
    r = CTCD::edgeEdgeCTCD(Eigen::Vector3d(0, 0, 0), Eigen::Vector3d(1, 0, 0), Eigen::Vector3d(2, 0, 0),
                           Eigen::Vector3d(2, 1, 0), Eigen::Vector3d(0, 0, 0), Eigen::Vector3d(1, 0, 0),
                           Eigen::Vector3d(2, 1, 0), Eigen::Vector3d(-1, -1, 0), 1e-6, t);
    out.precision(17);
    out << "ee_prim " << r << " " << t << "\n";
    r = CTCD::vertexFaceCTCD(Eigen::Vector3d(0.5, 0.5, 1), Eigen::Vector3d(0, 0, 0), Eigen::Vector3d(1, 0, 0),
                             Eigen::Vector3d(0, 1, 0), Eigen::Vector3d(0.5, 0.5, 1), Eigen::Vector3d(0, 0, 0),
                             Eigen::Vector3d(1, 0, 3), Eigen::Vector3d(0, 1, 0), 1e-6, t);
    out << "vf_prim " << r << " " << t << "\n";
    r = CTCD::vertexEdgeCTCD(Eigen::Vector3d(0.5, 0, 1), Eigen::Vector3d(0, 0, 0), Eigen::Vector3d(1, 0, 0),
                             Eigen::Vector3d(0.5, 0, -3), Eigen::Vector3d(0, 0, 0), Eigen::Vector3d(1, 0, 0), 1e-6, t);
    out << "ve_prim " << r << " " << t << "\n";
    r = CTCD::vertexVertexCTCD(Eigen::Vector3d(0, 0, 0), Eigen::Vector3d(1, 0, 0), Eigen::Vector3d(6, 0, 0),
                               Eigen::Vector3d(1, 0, 0), 1e-6, t);
    out << "vv_prim " << r << " " << t << "\n";
    return 0;
}
