// Umbrella header of the header-only template surface (the counterpart of
// gcurval/gcurval.h): plug-in vocabulary, the gfx950 engine template and the
// CPU engine template. Compile with hipcc --offload-arch=gfx950 -std=c++17
// and link libgcurval_amd.so (context, arena, copies through the C ABI); a
// host compiler sees the CPU engine only, as the reference's CPU build does
// (gcurval/arclength_calculate.h:1-30).
#pragma once

#include "base.h"
#if defined(__HIP__)
#include "arclength_calculate.h"
#endif
#include "arclength_calculate_cpu.h"
