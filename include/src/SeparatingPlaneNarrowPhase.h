// The reference's src/SeparatingPlaneNarrowPhase.h, at the path its example programs include ("../src/SeparatingPlaneNarrowPhase.h"): with
// -I<repo>/include/ccdamd that path resolves here.
#pragma once
#include "../ccdamd/SeparatingPlaneNarrowPhase.h"
