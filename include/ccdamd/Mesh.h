// Triangle mesh of the swept-CCD interface (reference: src/Mesh.h, src/Mesh.cpp).
#pragma once

#include "EigenTypes.h"

struct Mesh {
    Eigen::VectorXd vertices;
    Eigen::Matrix3Xi faces;   // one column per triangle

    // Mesh.cpp:13-20: faces that share any vertex.
    bool neighboringFaces(int face1, int face2) const
    {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                if (faces.coeff(i, face1) == faces.coeff(j, face2)) return true;
        return false;
    }
    // Mesh.cpp:5-11
    bool vertexOfFace(int vert, int face) const
    {
        return faces.coeff(0, face) == vert || faces.coeff(1, face) == vert || faces.coeff(2, face) == vert;
    }
};
