// Umbrella header of the header-only template surface (the counterpart of
// gcurval/gcurval.h): plug-in vocabulary plus the gfx950 engine template.
// Compile with hipcc --offload-arch=gfx950 -std=c++17 and link
// libgcurval_amd.so (context, arena, copies through the C ABI).
#pragma once

#include "base.h"
#include "arclength_calculate.h"
