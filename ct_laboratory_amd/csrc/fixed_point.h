// fixed_point.h -- device helpers shared by the LDS back projectors (3D bricks,
// 2D tiles): exact 64-bit fixed-point rounding of float contributions, the
// power-of-two scale that keeps a region's sums in range, and an IEEE float
// add to global memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctlab {

// Round-half-even of c (|c| < 2^50) to a 64-bit two's-complement integer: c
// is exact in double, and adding 1.5 * 2^52 lands in [2^52, 2^53), where the
// double spacing is 1, so that addition rounds c half-to-even and the integer
// sits in the low bits of the sum's representation.
__device__ __forceinline__ unsigned long long to_fixed(float c) {
    const double y = (double)c + 6755399441055744.0;  // 1.5 * 2^52
    return (unsigned long long)__double_as_longlong(y) - 0x4338000000000000ull;
}

// *p += v with the shader's IEEE single addition (compare-and-swap loop; each
// failed swap means another writer's add landed, so the loop ends).
__device__ __forceinline__ void add_f32_ieee(float* p, float v) {
    unsigned int* q = reinterpret_cast<unsigned int*>(p);
    unsigned int old = *q;
    while (true) {
        const unsigned int seen = atomicCAS(q, old, __float_as_uint(__uint_as_float(old) + v));
        if (seen == old) break;
        old = seen;
    }
}

// Scale 2^F for the fixed-point sums of one region (brick or tile).  cmax
// bounds a single ray's contribution to one cell (2 * max|value| * longest
// segment in a cell: the factor 2 covers the rounding of the length and of the
// product) and count the number of rays, so |sum| <= count * cmax per cell.
// 2^F maps that bound below 2^62 (no overflow of the signed 64-bit sums) and
// each single contribution below 2^50 (the range of to_fixed).  The scale is
// applied to each value as two float factors s1 * s2, so any F in [-250, 250]
// is exact.  fixed is false when the data or the map are not finite or a
// product could overflow a float: the caller then adds floats instead.
struct FixedScale {
    bool fixed;
    int F;
    float s1, s2;
};

__device__ __forceinline__ FixedScale fixed_scale(double cmax, double count, bool finite) {
    FixedScale fs;
    fs.fixed = finite && cmax <= 1.0e38;
    const double bound = cmax * (count > 0.0 ? count : 1.0);
    const int e = (bound > 0.0) ? ilogb(bound) + 1 : 0;  // bound < 2^e
    const int ec = (cmax > 0.0) ? ilogb(cmax) + 1 : 0;   // cmax < 2^ec
    fs.F = fs.fixed ? max(-250, min(min(62 - e, 50 - ec), 250)) : 0;
    fs.s1 = ldexpf(1.0f, fs.F / 2);
    fs.s2 = ldexpf(1.0f, fs.F - fs.F / 2);
    return fs;
}

}  // namespace ctlab
