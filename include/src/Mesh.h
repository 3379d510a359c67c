// The reference's src/Mesh.h, at the path its example programs include ("../src/Mesh.h"): with
// -I<repo>/include/ccdamd that path resolves here.
#pragma once
#include "../ccdamd/Mesh.h"
