// OBJ mesh input and output as the reference's drivers do it (loadMesh / writeMesh, defined in each of
// example/testVelocityFilter.cpp:11-77, AlecTest.cpp:16-62 and the other drivers).
//
// loadMesh reads the file one whitespace-separated character at a time: after a 'v' it reads three
// doubles, after an 'f' three 1-based vertex indices, and any other character is skipped on its own.
// So plain "v x y z" / "f a b c" files load as expected, while tokens such as "vn", "vt" or "f 1/2/3"
// stop the parse where the number fails to read, as in the reference.  A missing file leaves the
// outputs untouched.  writeMesh writes "v x y z " and "f a b c " lines at the stream's default precision
// (6 significant digits).
#pragma once

#include <fstream>
#include <vector>

#include "EigenTypes.h"

namespace ccdamd {

inline bool loadMesh(const char *filename, Eigen::VectorXd &verts, Eigen::Matrix3Xi &faces)
{
    std::ifstream in(filename);
    if (!in) return false;
    std::vector<double> pos;
    std::vector<int> idx;
    char c;
    while (in >> c) {
        if (c == 'v') {
            for (int i = 0; i < 3; i++) {
                double x;
                in >> x;
                pos.push_back(x);
            }
        }
        if (c == 'f') {
            for (int i = 0; i < 3; i++) {
                int x;
                in >> x;
                idx.push_back(x);
            }
        }
    }
    const std::ptrdiff_t nv = (std::ptrdiff_t)(pos.size() / 3), nf = (std::ptrdiff_t)(idx.size() / 3);
    verts.resize(3 * nv);
    for (std::ptrdiff_t i = 0; i < 3 * nv; i++) verts[i] = pos[i];
    faces.resize(3, nf);
    for (std::ptrdiff_t f = 0; f < nf; f++)
        for (int j = 0; j < 3; j++) faces.coeffRef(j, f) = idx[3 * f + j] - 1;
    return true;
}

inline bool writeMesh(const char *filename, const Eigen::VectorXd &verts, const Eigen::Matrix3Xi &faces)
{
    std::ofstream out(filename);
    if (!out) return false;
    for (std::ptrdiff_t i = 0; i < (std::ptrdiff_t)verts.size() / 3; i++) {
        out << "v ";
        for (int j = 0; j < 3; j++) out << verts[3 * i + j] << " ";
        out << std::endl;
    }
    for (std::ptrdiff_t f = 0; f < (std::ptrdiff_t)faces.cols(); f++) {
        out << "f ";
        for (int j = 0; j < 3; j++) out << faces.coeff(j, f) + 1 << " ";
        out << std::endl;
    }
    return true;
}

}   // namespace ccdamd
