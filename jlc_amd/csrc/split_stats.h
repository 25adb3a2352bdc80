// Per-entry statistics shared by the split forms' row / column statistics
// kernels (INT8 and BF16 slices, CRT moduli).
#pragma once

#include <math.h>

namespace jlc {

// Exponent marker for rows / columns holding Inf or NaN.
constexpr int STAT_NONFINITE = -100000;

// log2|v| for finite v > 0 from frexp and a float log2 of the mantissa.
__device__ __forceinline__ double stat_log2_abs(double v) {
    int fe;
    const double mnt = frexp(v, &fe);  // v = mnt * 2^fe, mnt in [0.5, 1)
    return (double)fe + (double)__log2f((float)mnt);
}

// Adds |x| = v to a row's running {max, sum of log2|x| over the finite
// non-zero entries, their count, any Inf/NaN}.
__device__ __forceinline__ void stat_add(double v, double& mx, double& slog, double& cnt, bool& bad) {
    bad |= !(v <= 1.7976931348623157e308);
    mx = fmax(mx, v);
    if (v > 0.0 && v <= 1.7976931348623157e308) {
        slog += stat_log2_abs(v);
        cnt += 1.0;
    }
}

}  // namespace jlc
