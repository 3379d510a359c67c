// The reference's src/KDOPBroadPhase.h, at the path its example programs include ("../src/KDOPBroadPhase.h"): with
// -I<repo>/include/ccdamd that path resolves here.
#pragma once
#include "../ccdamd/KDOPBroadPhase.h"
