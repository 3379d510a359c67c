// TEST INFRASTRUCTURE ONLY.  Exposes the reference's own Jenkins-Traub root finder with a C symbol so
// the oracle's tests can use it as a checker.  This file holds no reference code: it is compiled with
// -I pointing at the reference's src/ directory, in place, by oracle/Makefile, and the object it makes
// goes only to oracle/_ref/ (never into the product, never into git).
#include "rpoly.h"   // src/rpoly.h:49-261 (RootFinder::rpoly)

extern "C" int ref_rpoly(double *op, int degree, double *zeror, double *zeroi)
{
    RootFinder rf;   // CTCD.cpp:139-141 constructs one per call
    return rf.rpoly(op, degree, zeror, zeroi);
}
