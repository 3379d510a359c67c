// Reads an OBJ file with ccdamd::loadMesh, writes the vertices (float64) and faces (int32) as raw
// arrays, and writes the mesh back with ccdamd::writeMesh.
// Usage: objio_check in.obj verts.f64 faces.i32 out.obj
#include <fstream>
#include <iostream>

#include "ObjIO.h"

int main(int argc, char **argv)
{
    if (argc != 5) {
        std::cerr << "usage: objio_check in.obj verts.f64 faces.i32 out.obj" << std::endl;
        return 2;
    }
    Eigen::VectorXd v;
    Eigen::Matrix3Xi f;
    if (!ccdamd::loadMesh(argv[1], v, f)) return 1;
    std::ofstream vo(argv[2], std::ios::binary), fo(argv[3], std::ios::binary);
    for (std::ptrdiff_t i = 0; i < v.size(); i++) {
        double x = v[i];
        vo.write((const char *)&x, sizeof x);
    }
    for (std::ptrdiff_t c = 0; c < f.cols(); c++)
        for (int r = 0; r < 3; r++) {
            int x = f.coeff(r, c);
            fo.write((const char *)&x, sizeof x);
        }
    return ccdamd::writeMesh(argv[4], v, f) ? 0 : 1;
}
