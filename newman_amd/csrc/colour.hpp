// colour.hpp - the device colouring launcher (see colour.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"

namespace nm {

// Colours grid (nr x nc escape values, row-major, on the device) into out
// ((nr / sc) x (nc / sc) x 3 bytes, interleaved RGB, on the device) with the
// palette pal_rgba (4 bytes per entry, at least N entries, on the device).
// Returns 0 when the launch was accepted.
int colour_launch(hipStream_t stream, int num_cu, const EscapeValue* grid, int nr, int nc, int sc, int N,
                  const void* pal_rgba, int smooth, uint8_t* out);

}  // namespace nm
