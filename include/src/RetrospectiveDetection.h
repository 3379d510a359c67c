// The reference's src/RetrospectiveDetection.h, at the path its example programs include ("../src/RetrospectiveDetection.h"): with
// -I<repo>/include/ccdamd that path resolves here.
#pragma once
#include "../ccdamd/RetrospectiveDetection.h"
