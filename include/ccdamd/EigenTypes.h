// Vector and matrix types of the reference's interface (Eigen::Vector3d, VectorXd, Matrix3Xi).
// With Eigen on the include path these are Eigen's own; without it, EigenSubset.h defines the members
// the swept-CCD interface and the reference's example programs touch, so that this interface compiles
// and links on its own.
#pragma once

#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
#include <Eigen/Geometry>
#define CCDAMD_HAVE_EIGEN 1
#else
#include "EigenSubset.h"
#endif
