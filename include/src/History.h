// The reference's src/History.h, at the path its example programs include ("../src/History.h"): with
// -I<repo>/include/ccdamd that path resolves here.
#pragma once
#include "../ccdamd/History.h"
