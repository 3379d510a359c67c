// The reference's src/CTCDNarrowPhase.h, at the path its example programs include ("../src/CTCDNarrowPhase.h"): with
// -I<repo>/include/ccdamd that path resolves here.
#pragma once
#include "../ccdamd/CTCDNarrowPhase.h"
