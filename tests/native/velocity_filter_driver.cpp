// Test driver for the velocity filter's C++ interface, written the way the reference's
// example/testVelocityFilter.cpp uses the library: load two OBJ meshes, give the first numInfinite
// vertices infinite mass, VelocityFilter::velocityFilter, write the result.  Also exercises the Distance
// statics on one configuration.
// Usage: velocity_filter_driver start.obj end.obj numInfinite outerEta innerEta substep out.f64 out.obj out.txt
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "Distance.h"
#include "ObjIO.h"
#include "VelocityFilter.h"

int main(int argc, char **argv)
{
    if (argc != 10) {
        std::cerr << "usage: velocity_filter_driver start.obj end.obj numInfinite outerEta innerEta substep "
                     "out.f64 out.obj out.txt"
                  << std::endl;
        return 2;
    }
    Eigen::VectorXd q1, q2;
    Eigen::Matrix3Xi f1, f2;
    ccdamd::loadMesh(argv[1], q1, f1);
    ccdamd::loadMesh(argv[2], q2, f2);
    int numinfinite = (int)strtod(argv[3], nullptr);
    double outer = atof(argv[4]), inner = atof(argv[5]), substep = atof(argv[6]);
    std::cout << "Loaded " << q1.size() << " vertices, " << f1.size() << " faces." << std::endl;
    if (numinfinite >= q1.size() / 3) {
        std::cerr << "Bad value of numInfiniteMass" << std::endl;
        return 1;
    }
    Eigen::VectorXd invmasses(q1.size());
    for (int i = 0; i < 3 * numinfinite; i++) invmasses[i] = 0;
    for (int i = 3 * numinfinite; i < q1.size(); i++) invmasses[i] = 1.0;

    std::set<int> fixedVerts;
    for (int i = 0; i < numinfinite; i++) fixedVerts.insert(i);
    double closest = Distance::meshSelfDistance(q1, f1, fixedVerts);

    int ret = VelocityFilter::velocityFilter(q1, q2, f1, invmasses, outer, inner, .0001, substep);
    std::cout << ret << std::endl;
    ccdamd::writeMesh(argv[8], q2, f1);
    {
        std::ofstream bin(argv[7], std::ios::binary);
        for (int i = 0; i < q2.size(); i++) {
            double x = q2[i];
            bin.write((const char *)&x, sizeof x);
        }
    }

    std::ofstream out(argv[9]);
    out.precision(17);
    out << "result " << ret << "\n";
    out << "closest " << closest << "\n";
    // Distance statics on the first face and the vertex / edge after it.
    Eigen::Vector3d a(0, 0, 0), b(1, 0, 0), c(0, 1, 0), p(0.25, 0.25, 0.5), r(2, 2, 1), s(2, 3, 1);
    double b0, b1, b2, b3;
    Eigen::Vector3d vf = Distance::vertexFaceDistance(p, a, b, c, b0, b1, b2);
    out << "vf " << vf[0] << " " << vf[1] << " " << vf[2] << " " << b0 << " " << b1 << " " << b2 << "\n";
    Eigen::Vector3d ee = Distance::edgeEdgeDistance(a, b, r, s, b0, b1, b2, b3);
    out << "ee " << ee[0] << " " << ee[1] << " " << ee[2] << " " << b0 << " " << b1 << " " << b2 << " " << b3 << "\n";
    out << "plane " << Distance::vertexPlaneDistanceLessThan(p, a, b, c, 0.6) << " "
        << Distance::vertexPlaneDistanceLessThan(p, a, b, c, 0.4) << "\n";
    out << "line " << Distance::lineLineDistanceLessThan(a, b, r, s, 1.5) << " "
        << Distance::lineLineDistanceLessThan(a, b, r, s, 0.5) << "\n";
    return 0;
}
