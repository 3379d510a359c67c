// The reference's src/Stencils.h, at the path its example programs include ("../src/Stencils.h"): with
// -I<repo>/include/ccdamd that path resolves here.
#pragma once
#include "../ccdamd/Stencils.h"
