// ref_rng_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the reference's own particle generator, the `xorshf96` definition
// of src/services/init.cu:11-25 (its state words and the function), so the
// oracle's restatement of the generator can be checked against the original.
// init.cu itself needs CUDA, blitz++ and the missing tipsy reader; the
// Makefile cuts the generator's definition out of the file where it lies into
// oracle/_ref/ (git-ignored, never committed) and includes it here.
#include <cstdint>
#include <limits>

#include "xorshf96_from_init_cu.inc"

namespace {
bool g_saved = false;
unsigned long g_x0, g_y0, g_z0;
}  // namespace

extern "C" {

// n particles drawn as init.cu:47-53 draws them (for i: x, y, z), from the
// generator's initial state on every call.
void ref_generate(int64_t n, float *px, float *py, float *pz)
{
    if (!g_saved) {
        g_x0 = x;
        g_y0 = y;
        g_z0 = z;
        g_saved = true;
    }
    x = g_x0;
    y = g_y0;
    z = g_z0;
    for (int64_t i = 0; i < n; ++i) {
        px[i] = xorshf96();
        py[i] = xorshf96();
        pz[i] = xorshf96();
    }
}

}  // extern "C"
