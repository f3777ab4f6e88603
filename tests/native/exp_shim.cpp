// Host build of the device exp restatement (mesobasic_amd/csrc/lbm_exp.h),
// so a CPU test can compare it with the host libm bit for bit.
#include <math.h>
#include <stdint.h>

#include "lbm_exp.h"

extern "C" int64_t exp_shim_mismatches(const double* x, int64_t n, double* first_bad)
{
    int64_t bad = 0;
    for (int64_t i = 0; i < n; i++) {
        const double a = lbm::exp_libm(x[i]);
        const double b = exp(x[i]);
        if (__builtin_bit_cast(uint64_t, a) != __builtin_bit_cast(uint64_t, b)) {
            if (bad == 0 && first_bad) *first_bad = x[i];
            bad++;
        }
    }
    return bad;
}
