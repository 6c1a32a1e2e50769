// Host build of the device power function (gfm_fastmath.hpp) for its CPU
// test: compiled with g++ -ffp-contract=off, which rounds every operation as
// the device build does.
#include "gfm_fastmath.hpp"

extern "C" void pow_fast_n(int n, const double* x, const double* y, double* out) {
    for (int k = 0; k < n; ++k) out[k] = gfm::pow_fast(x[k], y[k]);
}
