// siddon_traverse.h -- one-ray-per-lane Siddon traversal for gfx950.
//
// The reference defines a ray's segments by enumerating every voxel-plane
// crossing, sorting the list and walking consecutive pairs
// (ct_laboratory/tomography/ct_projector_3d_torch.py:78-158 and :161-274;
// src/ct_projector_3d.cu:834-1012).  Here the per-axis crossing sequences are
// each monotone in the plane index, so their merge *is* the sorted list: the
// lane keeps one pending crossing per axis in registers and pops the
// smallest, with no scratch array and no sort.
//
// Bit-exactness.  Every crossing equals the reference's value bit for bit
// (torch: RN(RN(p-0.5-s)/dir); otf: RN(RN(p-0.5-s)*RN(1/dir))) and every
// segment's voxel is the reference's midpoint rule in the reference's
// operation order, so the voxel sequence equals the reference's.  The file is
// compiled with floating-point contraction off; fused multiply-adds appear only
// where written as __builtin_fmaf (the exact-division correction and the otf
// forms, which follow nvcc's default contraction of the CUDA source).
//
// Box clipping without losing exactness.  Only segments whose rounded
// midpoint falls inside the grid contribute.  The lane jumps each axis to its
// first crossing at or after t_lo = max(0, t_enter - tau) and stops after
// t_hi = min(1, t_exit + tau), where tau is four times a bound on the rounding
// error of the midpoint coordinate divided by |dir|.  Every crossing in
// [t_lo, t_hi] is produced exactly, and every segment of the full list that
// is not produced has its exact midpoint at least 2*E outside the box on the
// entry (or exit) axis, so its rounded index is out of range as well.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#pragma clang fp contract(off)

namespace ctlab {

enum : int { MODE_TORCH = 0, MODE_OTF = 1, MODE_MARCH = 2 };
enum : int { ZERO_THRESH = 0, ZERO_EXACT = 1 };

// Uniform per-launch geometry.  minv is row-major ND x ND.
template <int ND>
struct Grid {
    int n[ND];
    float minv[ND * ND];
    float b[ND];
    int diag;       // off-diagonal entries of minv are all zero
    int pow2;       // diag, and every diagonal entry of minv is +-2^k (normal)
    int finite;     // minv and b are finite
    int zero_rule;  // ZERO_THRESH: |dir| < 1e-12 has no crossings; ZERO_EXACT: dir == 0
};

template <int ND>
__host__ __device__ __forceinline__ Grid<ND> load_grid(const int* n, const float* __restrict__ minv,
                                                       const float* __restrict__ b, int zero_rule) {
    Grid<ND> g;
    int diag = 1, fin = 1;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        g.n[a] = n[a];
        g.b[a] = b[a];
        fin &= (fabsf(g.b[a]) <= 3.0e38f);
    }
#pragma unroll
    for (int i = 0; i < ND * ND; ++i) {
        g.minv[i] = minv[i];
        if ((i / ND) != (i % ND) && g.minv[i] != 0.0f) diag = 0;
        fin &= (fabsf(g.minv[i]) <= 3.0e38f);
    }
    int pow2 = diag;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const uint32_t u = __builtin_bit_cast(uint32_t, g.minv[a * ND + a]) & 0x7fffffffu;
        pow2 &= ((u & 0x007fffffu) == 0u && u >= 0x00800000u && u < 0x7f800000u);
    }
    g.diag = diag;
    g.pow2 = pow2;
    g.finite = fin;
    g.zero_rule = zero_rule;
    return g;
}

// Voxel window [lo, lo+ext) per axis.  Segments are produced only when their
// rounded voxel lies in the window, and the visitor receives the window index
// sum_a (i_a - lo_a) * stride_a (the last axis has stride 1).  The full grid
// is lo = 0, ext = n, stride = (n_y*n_z, n_z, 1): the reference's flat index.
// Strides must stay below 2^24.
template <int ND>
struct Window {
    int lo[ND];
    int ext[ND];
    int stride[ND];
};

template <int ND>
__host__ __device__ __forceinline__ Window<ND> full_window(const Grid<ND>& g) {
    Window<ND> w;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        w.lo[a] = 0;
        w.ext[a] = g.n[a];
    }
    w.stride[ND - 1] = 1;
#pragma unroll
    for (int a = ND - 2; a >= 0; --a) w.stride[a] = w.stride[a + 1] * g.n[a + 1];
    return w;
}

// Voxel coordinates of a ray end point, (p - b) @ M_inv^T (3d_torch.py:100-103;
// 2d_torch.py:26-29).  torch's CPU float32 matmul evaluates each 2- or 3-term
// dot product as the chain fma(x2, m2, fma(x1, m1, x0*m0)), and nvcc contracts
// the reference's apply_affine_inverse_3d (3d.cu:21-36) to the same chain, so
// both modes use it.  For a diagonal M_inv every order gives the same value.
template <int ND, int MODE>
__host__ __device__ __forceinline__ void to_voxel(const Grid<ND>& g, const float* p, float* out) {
    float sh[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) sh[j] = p[j] - g.b[j];
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        float acc = g.minv[a * ND] * sh[0];
#pragma unroll
        for (int j = 1; j < ND; ++j) acc = __builtin_fmaf(g.minv[a * ND + j], sh[j], acc);
        out[a] = acc;
    }
}

// to_voxel for a diagonal M_inv and a finite point: each off-diagonal term of
// the chain is a signed zero, and adding it changes at most the sign of a zero
// result.  No use of a voxel coordinate depends on the sign of a zero (it
// enters only as s in RN(f - s) with f = p - 0.5 != 0, in differences, and in
// magnitudes), so the traversal is the same.
template <int ND, int MODE, bool DIAG>
__host__ __device__ __forceinline__ void to_voxel_t(const Grid<ND>& g, const float* p, float* out) {
    if (DIAG) {
#pragma unroll
        for (int a = 0; a < ND; ++a) out[a] = g.minv[a * ND + a] * (p[a] - g.b[a]);
    } else {
        to_voxel<ND, MODE>(g, p, out);
    }
}

// Midpoint voxel coordinates from the world midpoint (3d_torch.py:244-250;
// 3d.cu:985-991).
template <int ND, int MODE, bool DIAG>
__host__ __device__ __forceinline__ void mid_coords(const Grid<ND>& g, const float* m, float* c) {
    float sh[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) sh[j] = m[j] - g.b[j];
    if (DIAG) {
        // Off-diagonal products are signed zeros; adding them leaves the value
        // (and its rounding to nearest-even) unchanged, so they are omitted.
#pragma unroll
        for (int a = 0; a < ND; ++a) c[a] = g.minv[a * ND + a] * sh[a];
    } else if (MODE == MODE_OTF) {
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            float acc = g.minv[a * ND] * sh[0];
#pragma unroll
            for (int j = 1; j < ND; ++j) acc = __builtin_fmaf(g.minv[a * ND + j], sh[j], acc);
            c[a] = acc;
        }
    } else {
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            float acc = g.minv[a * ND] * sh[0];
#pragma unroll
            for (int j = 1; j < ND; ++j) acc = acc + g.minv[a * ND + j] * sh[j];
            c[a] = acc;
        }
    }
}

// 24-bit multiply-add (v_mad_u32_u24 on the device).  Callers guarantee both
// factors are below 2^24 (the host rejects grids with n_y*n_z >= 2^24).
__host__ __device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b) + c;
#else
    return a * b + c;
#endif
}

// r * stride for an index r < 2^24 (the product is only used when r is in
// range): a shift or constant multiply when the stride is known at compile
// time, else the full-rate 24-bit multiply.
__host__ __device__ __forceinline__ uint32_t mul_stride(uint32_t r, int stride) {
    if (__builtin_constant_p(stride) && (stride & (stride - 1)) == 0) return r * (uint32_t)stride;
    return mad24(r, (uint32_t)stride, 0u);
}

// Round half to even, range check and flat index ((i*n_y + j)*n_z + k as
// i*(n_y*n_z) + (j*n_z + k)).  A NaN coordinate clamps to -1 (fmaxf returns the
// non-NaN operand) and is out of range, as round(NaN).long() is in the
// reference.
template <int ND>
__host__ __device__ __forceinline__ bool round_voxel(const Grid<ND>& g, const float* c, int32_t* r) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float f = __builtin_rintf(c[a]);  // v_rndne_f32: ties to even
        // Clamp into int range first (the host conversion of out-of-range
        // floats is undefined); every clamped value is out of the grid anyway.
        const float fc = fminf(fmaxf(f, -1.0f), 2147483520.0f);
        r[a] = (int32_t)fc;
        ok = ok && ((uint32_t)r[a] < (uint32_t)g.n[a]);
    }
    return ok;
}

template <int ND>
__host__ __device__ __forceinline__ bool flat_voxel(const Grid<ND>& g, const float* c, uint32_t* idx) {
    int32_t r[ND];
    const bool ok = round_voxel<ND>(g, c, r);
    if (ND == 3) {
        *idx = mad24((uint32_t)r[0], (uint32_t)(g.n[1] * g.n[2]), mad24((uint32_t)r[1], (uint32_t)g.n[2],
                                                                       (uint32_t)r[ND - 1]));
    } else {
        *idx = mad24((uint32_t)r[0], (uint32_t)g.n[ND - 1], (uint32_t)r[ND - 1]);
    }
    return ok;
}

// Same as flat_voxel for finite coordinates with |c| < 2^22, in two integer-
// exact steps per axis: RN(c + 1.5*2^23) lies in [2^23, 2^24) where the float
// spacing is 1, so the addition itself rounds c half-to-even, and the bit
// pattern minus that of 1.5*2^23 is the rounded integer.  kb[a] holds that
// bit pattern plus the window's lower corner, so one subtraction gives the
// window index.  The clipped
// traversal only evaluates midpoints within a few voxels of the grid (and the
// grid dimensions are below 2^24 with E < 1), so |c| stays far below 2^22.
//
// BIASED: c already holds RN(c + 1.5*2^23) (MidIndex::from_ends_biased).
template <int ND, bool BIASED = false>
__host__ __device__ __forceinline__ bool flat_voxel_bounded(const Window<ND>& w, const uint32_t* kb, const float* c,
                                                            uint32_t* idx) {
    const float kMagic = 12582912.0f;  // 1.5 * 2^23
    uint32_t r[ND];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float f = BIASED ? c[a] : c[a] + kMagic;
        r[a] = __builtin_bit_cast(uint32_t, f) - kb[a];
        ok = ok && (r[a] < (uint32_t)w.ext[a]);
    }
    // nested multiply-adds: one full-rate v_mad_u32_u24 (or shift-add) per axis
    if (ND == 3) {
        *idx = mad24(r[0], (uint32_t)w.stride[0], mul_stride(r[1], w.stride[1]) + r[ND - 1]);
    } else {
        *idx = mad24(r[0], (uint32_t)w.stride[0], r[ND - 1]);
    }
    return ok;
}

// Running sum of squared coordinate differences of a segment: the torch
// rule's RN(RN(ss) + RN(d*d)) on the host (test harness, bit for bit with the
// oracle), one fma on the device (at most one rounding per term either way;
// lengths only scale contributions, voxel indices never depend on them).
__host__ __device__ __forceinline__ float add_sq(float ss, float d, bool first) {
#if defined(__HIP_DEVICE_COMPILE__)
    return first ? d * d : __builtin_fmaf(d, d, ss);
#else
    const float sq = d * d;
    return first ? sq : ss + sq;
#endif
}

// Square root of a segment length.  The device uses the hardware
// v_sqrt_f32 (within 1 ulp); the host build (test harness) uses the correctly
// rounded sqrtf.  Lengths only scale contributions; voxel indices never depend
// on them.
__host__ __device__ __forceinline__ float seg_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return __builtin_sqrtf(x);
#endif
}

// Plane crossing with plane coordinate f = p - 0.5 (exact), voxel start s,
// voxel direction d and y = RN(1/d):
//   torch: RN(RN(f-s)/d), computed as q0 = RN(num*y); r = num - q0*d (exact,
//          one fma); RN(q0 + r*y).  With y = RN(1/d) and no over/underflow this
//          is the correctly rounded quotient (Markstein's theorem; also checked
//          against IEEE division on 3.5e8 random and adversarial operand
//          pairs), i.e. bit-identical to the reference's division.
//   otf:   RN(RN(f-s)*y), the reference's reciprocal form (3d.cu:866-868).
template <int MODE>
__host__ __device__ __forceinline__ float crossing_f(float f, float s, float d, float y) {
    const float num = f - s;
    const float q0 = num * y;
    if (MODE == MODE_OTF) return q0;
    const float r = __builtin_fmaf(-q0, d, num);
    return __builtin_fmaf(r, y, q0);
}

template <int MODE>
__host__ __device__ __forceinline__ float crossing_p(int p, float s, float d, float y) {
    return crossing_f<MODE>((float)p - 0.5f, s, d, y);
}

// First plane index p in [plo, phi] (in the axis' traversal direction) whose
// crossing is >= t_lo, or one past the range if none.  Crossings are monotone
// in p, so the guess from the coordinate at t_lo is corrected by at most a few
// steps; the clamps keep a NaN guess inside [plo-1, phi+1] so both loops stay
// bounded.
template <int MODE>
__host__ __device__ __forceinline__ int first_plane(float s, float d, float y, int plo, int phi, int step,
                                                    float t_lo) {
    const float c = s + t_lo * d;
    int p;
    if (step > 0) {
        const float f = ceilf(c + 0.5f);
        p = (f >= (float)plo) ? ((f <= (float)(phi + 1)) ? (int)f : phi + 1) : plo;
        // The guess is off by at most a step or two: no unrolling.
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
        while (p > plo && crossing_p<MODE>(p - 1, s, d, y) >= t_lo) --p;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
        while (p <= phi && crossing_p<MODE>(p, s, d, y) < t_lo) ++p;
    } else {
        const float f = floorf(c + 0.5f);
        p = (f >= (float)(plo - 1)) ? ((f <= (float)phi) ? (int)f : phi) : plo - 1;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
        while (p < phi && crossing_p<MODE>(p + 1, s, d, y) >= t_lo) ++p;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
        while (p >= plo && crossing_p<MODE>(p, s, d, y) < t_lo) --p;
    }
    return p;
}

// Merge of the per-axis crossing sequences, held in scalar registers (no
// arrays, so nothing is dynamically indexed).  Axis a keeps the plane
// coordinate f_a = p_a - 0.5 of its pending crossing, e_a = the plane
// coordinate one past its last plane, and the pending crossing value n_a (INF
// when exhausted or past t = 1).  An axis traversed downwards is stored
// mirrored (s, d, y, f and e negated): every crossing value is unchanged, since
// RN(-x) = -RN(x) and the fma forms see the same products, so each axis steps
// by +1 and needs no step register.
template <int ND, int MODE>
struct Merge {
    float s0 = 0, s1 = 0, s2 = 0, d0 = 1, d1 = 1, d2 = 1, y0 = 1, y1 = 1, y2 = 1;
    float f0 = 0, f1 = 0, f2 = 0, e0 = 0, e1 = 0, e2 = 0;
    float n0 = __builtin_huge_valf(), n1 = __builtin_huge_valf(), n2 = __builtin_huge_valf();

    // Initialises axis a (a is a compile-time constant after unrolling).
    __host__ __device__ __forceinline__ void set(int a, float s, float d, float y, int p, int step, int plo,
                                                 int phi, float t_next) {
        float f = (float)p - 0.5f;
        float e = (step > 0) ? (float)phi + 0.5f : (float)plo - 1.5f;
        if (step < 0) {
            s = -s;
            d = -d;
            y = -y;
            f = -f;
            e = -e;
        }
        if (a == 0) { s0 = s; d0 = d; y0 = y; f0 = f; e0 = e; n0 = t_next; }
        if (a == 1) { s1 = s; d1 = d; y1 = y; f1 = f; e1 = e; n1 = t_next; }
        if (a == 2) { s2 = s; d2 = d; y2 = y; f2 = f; e2 = e; n2 = t_next; }
    }

    // Advanced candidate of one axis: the crossing after its pending one.
    // EXH: past its last plane the axis is exhausted (INF).  Without EXH the
    // caller stops before any crossing beyond the last plane matters.  Values
    // above 1 need no test: pop() only accepts crossings <= t_hi <= 1.
    template <bool EXH>
    __host__ __device__ __forceinline__ static float advance(float f, float s, float d, float y, float e,
                                                             float* f_out) {
        const float fn = f + 1.0f;
        const float tn = crossing_f<MODE>(fn, s, d, y);
        *f_out = fn;
        if (!EXH) return tn;
        return (fn == e) ? __builtin_huge_valf() : tn;
    }

    // Pops the smallest pending crossing if it is <= t_hi.  Every axis
    // evaluates its own candidate from its own registers and only the results
    // are selected: selecting the operands instead lets the compiler turn the
    // per-axis state into an indexed stack array (scratch memory).
    //
    // TIES: every axis whose pending crossing equals the popped value
    // advances at once, so a crossing shared by several axes comes out once
    // (the segment walks, whose zero-length pieces between equal crossings
    // contribute nothing).  Otherwise one axis advances per pop and equal
    // crossings come out once per axis (the crossing lists).
    template <bool EXH = true, bool TIES = false>
    __host__ __device__ __forceinline__ bool pop(float t_hi, float& t) {
        const float m = (ND == 3) ? fminf(fminf(n0, n1), n2) : fminf(n0, n1);
        if (!(m <= t_hi)) return false;
        const bool k0 = (n0 == m);
        const bool k1 = (TIES || !k0) && (n1 == m);
        float g0, g1, g2 = 0.0f;
        const float c0 = advance<EXH>(f0, s0, d0, y0, e0, &g0);
        const float c1 = advance<EXH>(f1, s1, d1, y1, e1, &g1);
        f0 = k0 ? g0 : f0;
        n0 = k0 ? c0 : n0;
        if (ND == 3) {
            const bool k2 = TIES ? (n2 == m) : (!k0 && !k1);
            const float c2 = advance<EXH>(f2, s2, d2, y2, e2, &g2);
            f1 = k1 ? g1 : f1;
            n1 = k1 ? c1 : n1;
            f2 = k2 ? g2 : f2;
            n2 = k2 ? c2 : n2;
        } else if (TIES) {
            f1 = k1 ? g1 : f1;
            n1 = k1 ? c1 : n1;
        } else {
            f1 = k0 ? f1 : g1;
            n1 = k0 ? n1 : c1;
        }
        t = m;
        return true;
    }
};

// 0.5f * x and 2.0f * x (round to nearest even, denormals and INF/NaN
// included) in integer operations on the bit pattern.  For wave-uniform x they
// run on the scalar unit and stay in scalar registers, where the float
// products would occupy (or be recomputed in) vector registers: gfx950 has no
// scalar float arithmetic.
__host__ __device__ __forceinline__ float half_exact(float x) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    const uint32_t sign = u & 0x80000000u, mag = u & 0x7fffffffu;
    uint32_t r;
    if (mag >= 0x7f800000u) {
        r = u;  // INF, NaN
    } else if (mag >= 0x01000000u) {
        r = u - 0x00800000u;  // exponent >= 2: the result is normal and exact
    } else {
        // exponent 0 or 1: value = mag * 2^-149, so halve mag with ties to even
        r = sign | ((mag >> 1) + (mag & (mag >> 1) & 1u));
    }
    return __builtin_bit_cast(float, r);
}

__host__ __device__ __forceinline__ float twice_exact(float x) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    const uint32_t sign = u & 0x80000000u, mag = u & 0x7fffffffu;
    uint32_t r;
    if (mag >= 0x7f800000u) {
        r = u;  // INF, NaN
    } else if (mag >= 0x7f000000u) {
        r = sign | 0x7f800000u;  // exponent 254: overflows to INF
    } else if (mag >= 0x00800000u) {
        r = u + 0x00800000u;  // normal
    } else {
        r = sign | (mag << 1);  // zero or denormal: exact, possibly normal
    }
    return __builtin_bit_cast(float, r);
}

// Midpoint voxel of a segment from its end points, folded: with S = P0+P1,
// RN(Minv*RN(RN(0.5*S) - b)) == RN((0.5*Minv) * RN(S - 2b)) exactly (scaling
// by a power of two commutes with rounding), so h = 0.5*Minv and b2 = 2b are
// precomputed.  Diagonal grids use only h[a][a]: the off-diagonal products are
// signed zeros and adding them changes neither value nor rounding.
template <int ND, bool DIAG>
struct MidIndex {
    float h[ND * ND];
    float b2[ND];
    __host__ __device__ __forceinline__ void init(const Grid<ND>& g) {
#pragma unroll
        for (int i = 0; i < ND * ND; ++i) h[i] = half_exact(g.minv[i]);
#pragma unroll
        for (int a = 0; a < ND; ++a) b2[a] = twice_exact(g.b[a]);
    }
    // Diagonal grid whose h[a][a] are powers of two (Grid::pow2): returns
    // RN(c + 1.5*2^23) for flat_voxel_bounded<ND, true>, as one fma per axis.
    // h * sh is exact (a power-of-two scaling that neither overflows, since
    // |c| < 2^22, nor rounds unless |h * sh| < 2^-126), so
    // fma(h, sh, 1.5*2^23) = RN(h*sh + 1.5*2^23) = RN(RN(h*sh) + 1.5*2^23);
    // below 2^-126 both sides are 1.5*2^23.
    __host__ __device__ __forceinline__ void from_ends_biased(const float* p0, const float* p1, float* c) const {
#pragma unroll
        for (int a = 0; a < ND; ++a) c[a] = __builtin_fmaf(h[a * ND + a], (p0[a] + p1[a]) - b2[a], 12582912.0f);
    }
    __host__ __device__ __forceinline__ void from_ends(const float* p0, const float* p1, float* c) const {
        float sh[ND];
#pragma unroll
        for (int a = 0; a < ND; ++a) sh[a] = (p0[a] + p1[a]) - b2[a];
        if (DIAG) {
#pragma unroll
            for (int a = 0; a < ND; ++a) c[a] = h[a * ND + a] * sh[a];
        } else {
#pragma unroll
            for (int a = 0; a < ND; ++a) {
                float acc = h[a * ND] * sh[0];
#pragma unroll
                for (int j = 1; j < ND; ++j) acc = acc + h[a * ND + j] * sh[j];
                c[a] = acc;
            }
        }
    }
};

// Visits every segment of one ray whose rounded midpoint lies in the window,
// in increasing t, as visit(window_flat_index, length).  Zero-length segments
// are not visited (they contribute exactly zero to both projections) unless
// ZERO is set.  The
// window's bounding planes take the place of the grid's in the clipping
// argument above, so each segment of the full traversal is visited by exactly
// the window that contains its voxel.
//
// PAIR unrolls the segment loop by two (see walk below).
//
// POW2 (callers pass it only for a grid with Grid::pow2 set) forms the torch
// rule's midpoint index with MidIndex::from_ends_biased: the same indices in
// fewer instructions.  ZERO also visits zero-length segments, for callers
// whose visit adds an exact zero for them (finite values).
//
// e_bound >= 0 replaces the per-ray midpoint error bound E by a caller's bound
// that holds for this ray on every axis (e.g. the maximum over a view): a
// larger E only widens [t_lo, t_hi], which produces the same segments.
template <int ND, int MODE, bool DIAG, bool PAIR = true, bool POW2 = false, bool ZERO = false, class Visit>
__host__ __device__ __forceinline__ void trace_ray_window(const Grid<ND>& g, const Window<ND>& win,
                                                          const float* src, const float* dst, Visit&& visit,
                                                          float e_bound = -1.0f) {
    static_assert(!POW2 || DIAG, "POW2 grids are diagonal");
    constexpr bool kFold = POW2 && MODE == MODE_TORCH;
    if (!g.finite) return;  // the reference's indices are all out of range then
    // Non-finite endpoints make every reference world point INF or NaN, and
    // every such segment is masked (3d_torch.py:233-260): no contribution.
    bool fin = true;
#pragma unroll
    for (int a = 0; a < ND; ++a) fin = fin && (fabsf(src[a]) <= 3.0e38f) && (fabsf(dst[a]) <= 3.0e38f);
    if (!fin) return;
    float vw[ND], sv[ND], dv[ND];
#pragma unroll
    for (int a = 0; a < ND; ++a) vw[a] = dst[a] - src[a];

    float rl = 0.0f;
    if (MODE == MODE_OTF) {
        float acc = vw[0] * vw[0];
#pragma unroll
        for (int a = 1; a < ND; ++a) acc = __builtin_fmaf(vw[a], vw[a], acc);
        rl = __builtin_sqrtf(acc);
        if (rl < 1.0e-12f) return;
    }

    to_voxel_t<ND, MODE, DIAG>(g, src, sv);
    to_voxel_t<ND, MODE, DIAG>(g, dst, dv);

    float dir[ND], y[ND];
    int step[ND];
    float t_lo = 0.0f, t_hi = 1.0f;
    const float kInf = __builtin_huge_valf();
    float e_max = 0.0f, t_cap = kInf;  // t_cap: first crossing of a plane beyond the window

#pragma unroll
    for (int a = 0; a < ND; ++a) {
        dir[a] = dv[a] - sv[a];
        // Rounding-error bound of a midpoint coordinate on this axis, voxel units.
        float E = e_bound;
        if (!(E >= 0.0f)) {
            float mag = fabsf(sv[a]) + fabsf(dv[a]) + (float)g.n[a] + 1.0f;
#pragma unroll
            for (int j = 0; j < ND; ++j)
                mag += fabsf(g.minv[a * ND + j]) * (fabsf(src[j]) + fabsf(dst[j]) + fabsf(g.b[j]));
            E = mag * 7.62939453125e-06f;  // 2^-17
        }
        // Midpoint coordinates stay within n + 5E of the grid; with E <= 8 they
        // are far below the 2^22 bound of flat_voxel_bounded.  Endpoints that far
        // out (~1e11 voxels) carry no usable float32 crossings anyway.
        if (!(E <= 8.0f)) return;
        e_max = fmaxf(e_max, E);
        bool active;
        if (MODE == MODE_OTF) {
            active = fabsf(dir[a]) > 1.0e-12f;
        } else {
            active = (g.zero_rule == ZERO_EXACT) ? (dir[a] != 0.0f) : !(fabsf(dir[a]) < 1.0e-12f);
        }
        step[a] = 0;
        y[a] = 1.0f;
        if (!active) {
            // Constant coordinate: a ray outside the slab by more than E has no
            // in-range segment at all.
            if (sv[a] < (float)win.lo[a] - 0.5f - E || sv[a] > (float)(win.lo[a] + win.ext[a]) - 0.5f + E) return;
            continue;
        }
        y[a] = 1.0f / dir[a];
        const bool up = dir[a] > 0.0f;
        step[a] = up ? 1 : -1;
        const int plo = win.lo[a], phi = win.lo[a] + win.ext[a];
        const float t_in = crossing_p<MODE>(up ? plo : phi, sv[a], dir[a], y[a]);
        const float t_out = crossing_p<MODE>(up ? phi : plo, sv[a], dir[a], y[a]);
        // 4E/|dir| from y = RN(1/dir): within 2^-23 relative, far inside the
        // factor-of-two slack of the clipping argument.
        const float tau = 4.0f * E * fabsf(y[a]) + 9.5367431640625e-07f;  // + 2^-20
        t_lo = fmaxf(t_lo, t_in - tau);
        t_hi = fminf(t_hi, t_out + tau);
        t_cap = fminf(t_cap, crossing_p<MODE>(up ? phi + 1 : plo - 1, sv[a], dir[a], y[a]));
    }
    if (!(t_lo <= t_hi)) return;

    // Jump each axis to its first crossing >= t_lo.
    Merge<ND, MODE> mg;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        if (step[a] == 0) continue;
        const int plo = win.lo[a], phi = win.lo[a] + win.ext[a];
        const int p = first_plane<MODE>(sv[a], dir[a], y[a], plo, phi, step[a], t_lo);
        float t = kInf;
        if (p >= plo && p <= phi) {
            t = crossing_p<MODE>(p, sv[a], dir[a], y[a]);
            if (t > 1.0f) t = kInf;
        }
        mg.set(a, sv[a], dir[a], y[a], p, step[a], plo, phi, t);
    }

    // Two forms of the segment loop.  With every midpoint error bound below
    // one half (all practical geometries) the walk stops at the first crossing
    // of a plane beyond the window, t_cap, and needs no per-axis exhaustion
    // test: a crossing of plane phi + 1 is a reference crossing when the
    // window ends inside the grid, and otherwise (a plane past the grid) it
    // ends only a segment at least half a voxel beyond the grid, which no
    // rounding within E < 1/2 brings back.  Segments past that plane are a
    // voxel or more outside the window.  Larger bounds take the exhaustive
    // form.
    MidIndex<ND, DIAG> mi;
    mi.init(g);
    // bit pattern of 1.5 * 2^23 + lo per axis, read from the first active lane:
    // every caller passes one window per wave (the whole grid, a block's tile
    // or brick), so it sits in scalar registers
    uint32_t kb[ND];
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        kb[a] = 0x4b400000u + (uint32_t)win.lo[a];
#if defined(__HIP_DEVICE_COMPILE__)
        kb[a] = __builtin_amdgcn_readfirstlane(kb[a]);
#endif
    }
    auto walk = [&](auto exh, float t_stop) {
        constexpr bool EXH = decltype(exh)::value;
        float t0;
        if (MODE == MODE_OTF && t_lo == 0.0f) {
            t0 = 0.0f;  // endpoint inserted by collect_hits (3d.cu:890)
        } else if (!mg.template pop<EXH, true>(t_stop, t0)) {
            return;
        }
        const bool close_at_one = (MODE == MODE_OTF) && (t_stop == 1.0f);
        // Next segment end: the next crossing, or t = 1 closing an otf ray.
        auto next = [&](float& t1, bool& last) -> bool {
            if (mg.template pop<EXH, true>(t_stop, t1)) return true;
            if (close_at_one && t0 < 1.0f) {
                t1 = 1.0f;  // endpoint inserted by collect_hits (3d.cu:891)
                last = true;
                return true;
            }
            return false;
        };
        // Segment [t0, t1]; pa holds P(t0) and pb receives P(t1) (torch rule).
        auto segment = [&](const float* pa, float* pb, float t1) {
            float c[ND];
            float len;
            if (MODE == MODE_TORCH) {
                float ss = 0.0f;
#pragma unroll
                for (int a = 0; a < ND; ++a) {
                    pb[a] = src[a] + t1 * vw[a];
                    ss = add_sq(ss, pb[a] - pa[a], a == 0);
                }
                len = seg_sqrt(ss);
                if (kFold)
                    mi.from_ends_biased(pa, pb, c);
                else
                    mi.from_ends(pa, pb, c);
            } else {
                float m[ND];
                const float tm = 0.5f * (t0 + t1);
#pragma unroll
                for (int a = 0; a < ND; ++a) m[a] = __builtin_fmaf(tm, vw[a], src[a]);
                len = (t1 > t0) ? (t1 - t0) * rl : 0.0f;
                mid_coords<ND, MODE, DIAG>(g, m, c);
            }
            t0 = t1;
            uint32_t idx;
            const bool in = flat_voxel_bounded<ND, kFold>(win, kb, c, &idx);
            if (in && (ZERO || len > 0.0f)) visit(idx, len);
        };
        // PAIR: two segments per trip with the end-point registers swapping
        // roles, so no point is copied (fewer instructions, more registers).
        float p0[ND], p1[ND];
#pragma unroll
        for (int a = 0; a < ND; ++a) p0[a] = src[a] + t0 * vw[a];
        while (true) {
            float t1;
            bool last = false;
            if (!next(t1, last)) break;
            segment(p0, p1, t1);
            if (last) break;
            if (!PAIR) {
#pragma unroll
                for (int a = 0; a < ND; ++a) p0[a] = p1[a];
                continue;
            }
            if (!next(t1, last)) break;
            segment(p1, p0, t1);
            if (last) break;
        }
    };
    if (e_max < 0.5f)
        walk(std::integral_constant<bool, false>(), fminf(t_hi, t_cap));
    else
        walk(std::integral_constant<bool, true>(), t_hi);
}

// The whole grid: visit(flat_index, length) with the reference's flat index.
template <int ND, int MODE, bool DIAG, bool POW2 = false, class Visit>
__host__ __device__ __forceinline__ void trace_ray(const Grid<ND>& g, const float* src, const float* dst,
                                                   Visit&& visit) {
    trace_ray_window<ND, MODE, DIAG, true, POW2>(g, full_window(g), src, dst, visit);
}

// Emits, in increasing order, every plane crossing t in [0, 1] of the whole
// ray: the sorted list of compute_intersections_{2d,3d}
// (ct_projector_3d_torch.py:78-158, src/ct_projector_3d.cu:42-96) without its
// INF padding.  Crossings use the division form RN(RN(p-0.5-s)/dir); an axis
// has none when the zero rule says so.  Returns the number emitted.
template <int ND, class Emit>
__host__ __device__ __forceinline__ int emit_crossings(const Grid<ND>& g, const float* src, const float* dst,
                                                       Emit&& emit) {
    float sv[ND], dv[ND];
    to_voxel<ND, MODE_TORCH>(g, src, sv);
    to_voxel<ND, MODE_TORCH>(g, dst, dv);
    Merge<ND, MODE_TORCH> mg;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float dir = dv[a] - sv[a];
        const bool active = (g.zero_rule == ZERO_EXACT) ? (dir != 0.0f) : !(fabsf(dir) < 1.0e-12f);
        if (!active) continue;
        const float y = 1.0f / dir;
        const int step = dir > 0.0f ? 1 : -1;
        const int p = first_plane<MODE_TORCH>(sv[a], dir, y, 0, g.n[a], step, 0.0f);
        float t = __builtin_huge_valf();
        if (p >= 0 && p <= g.n[a]) {
            t = crossing_p<MODE_TORCH>(p, sv[a], dir, y);
            if (t > 1.0f) t = __builtin_huge_valf();
        }
        mg.set(a, sv[a], dir, y, p, step, 0, g.n[a], t);
    }
    int count = 0;
    float t;
    while (mg.pop(1.0f, t)) {
        emit(t);
        ++count;
    }
    return count;
}

// ---------------------------------------------------------------------------
// The reference's 2D on-the-fly kernels (MODE_MARCH)
// ---------------------------------------------------------------------------

// Float to int the way the GPU's conversion instructions do it (NaN gives 0,
// large magnitudes saturate), except that the saturation is at +-2^30 rather
// than the int limits: such an index is far outside any grid either way, and
// stays outside (and free of overflow) over the march's bounded step count.
// (A plain C cast is undefined for these inputs.)
__host__ __device__ __forceinline__ int cvt_sat_i32(float x) {
    if (!(x == x)) return 0;
    if (x >= 1073741824.0f) return 1073741824;
    if (x <= -1073741824.0f) return -1073741824;
    return (int)x;
}

// |(vx, vy)|: the square root of the double sum of squares, rounded to float.
// Host and device compute it with the same correctly rounded double operations.
__host__ __device__ __forceinline__ float hypot_f32(float vx, float vy) {
    const double s = __builtin_fma((double)vx, (double)vx, (double)vy * (double)vy);
    return (float)__builtin_sqrt(s);
}

// Visits the steps of the incremental Siddon march of
// forward/back_project_2d_on_the_fly_kernel (src/ct_projector_2d.cu:246-331,
// 334-413), in order, as visit(flat_index, seg) for every step whose (row,
// column) lies in the grid, including steps of zero or negative length: the
// entry voxel is taken by floor (ceil - 1 going down) of the entry point's
// coordinate, while voxel r spans [r - 1/2, r + 1/2], so an entry coordinate
// with a fractional part of 1/2 or more starts one voxel early; the first
// crossing then lies behind the entry point, and its step has negative length
// (t accumulated: t_next += |1/v|, seg = (t_hit - t_cur) * |dst - src|).
// Every operation follows the CUDA source as nvcc compiles it by default: the
// forms a*b + c are fused (the affine map, the entry point, the first
// crossing), the rest rounded separately.  The ray length is hypot_f32 (the
// reference calls CUDA's hypotf, a different rounding of the same quantity).
// The step count is capped at 2 (n_row + n_col) + 8, which finite rays never
// reach (each step moves one index by one across the box, plus the entry
// correction); the cap only bounds the loop for non-finite input.
template <class Visit>
__host__ __device__ __forceinline__ void march_ray_2d(const Grid<2>& g, const float* src, const float* dst,
                                                      Visit&& visit) {
    const float vx = dst[0] - src[0], vy = dst[1] - src[1];
    const float ray_len = hypot_f32(vx, vy);
    if (ray_len < 1e-12f) return;
    float sv[2], dv[2];
    to_voxel<2, MODE_OTF>(g, src, sv);
    to_voxel<2, MODE_OTF>(g, dst, dv);
    const float sr = sv[0], sc = sv[1];
    const float vr = dv[0] - sr, vc = dv[1] - sc;
    const float inv_vr = (fabsf(vr) < 1e-12f) ? 1e32f : 1.0f / vr;
    const float inv_vc = (fabsf(vc) < 1e-12f) ? 1e32f : 1.0f / vc;
    const int n_row = g.n[0], n_col = g.n[1];
    const float t_r0 = (-0.5f - sr) * inv_vr;
    const float t_r1 = (((float)n_row - 0.5f) - sr) * inv_vr;
    const float t_c0 = (-0.5f - sc) * inv_vc;
    const float t_c1 = (((float)n_col - 0.5f) - sc) * inv_vc;
    float t_enter = fmaxf(fminf(t_r0, t_r1), fminf(t_c0, t_c1));
    float t_exit = fminf(fmaxf(t_r0, t_r1), fmaxf(t_c0, t_c1));
    if (t_exit <= t_enter) return;
    t_enter = fmaxf(t_enter, 0.0f);
    t_exit = fminf(t_exit, 1.0f);
    const float r_enter = __builtin_fmaf(t_enter, vr, sr);
    const float c_enter = __builtin_fmaf(t_enter, vc, sc);
    const bool up_r = vr >= 0.0f, up_c = vc >= 0.0f;
    int ri = up_r ? cvt_sat_i32(floorf(r_enter)) : cvt_sat_i32(ceilf(r_enter)) - 1;
    int ci = up_c ? cvt_sat_i32(floorf(c_enter)) : cvt_sat_i32(ceilf(c_enter)) - 1;
    float t_next_r = up_r ? __builtin_fmaf(((float)ri + 0.5f) - r_enter, inv_vr, t_enter)
                          : __builtin_fmaf(((float)ri - 0.5f) - r_enter, inv_vr, t_enter);
    float t_next_c = up_c ? __builtin_fmaf(((float)ci + 0.5f) - c_enter, inv_vc, t_enter)
                          : __builtin_fmaf(((float)ci - 0.5f) - c_enter, inv_vc, t_enter);
    const float dt_r = fabsf(inv_vr), dt_c = fabsf(inv_vc);
    const int step_r = up_r ? 1 : -1, step_c = up_c ? 1 : -1;
    float t_cur = t_enter;
    const int cap = 2 * (n_row + n_col) + 8;
    for (int k = 0; k < cap && t_cur < t_exit; ++k) {
        const bool step_row = t_next_r < t_next_c;
        const float t_hit = fminf(step_row ? t_next_r : t_next_c, t_exit);
        if (ri >= 0 && ri < n_row && ci >= 0 && ci < n_col)
            visit((uint32_t)ri * (uint32_t)n_col + (uint32_t)ci, (t_hit - t_cur) * ray_len);
        t_cur = t_hit;
        if (step_row) {
            t_next_r += dt_r;
            ri += step_r;
        } else {
            t_next_c += dt_c;
            ci += step_c;
        }
    }
}

// Segment geometry of one pair of crossings given as t values, in the op order
// of the torch consumer (3d_torch.py:219-271): world points src + t (dst-src),
// length |P1 - P0|, midpoint voxel by round-half-even.  p0 holds P(t0) on
// entry and P(t1) on return.  Returns whether the voxel is in the grid; *idx
// and *len receive the flat voxel index and the length.
//
// segment_mid is the same without the voxel lookup: it returns the length and
// leaves the midpoint's continuous voxel coordinates in c.
template <int ND, bool DIAG>
__host__ __device__ __forceinline__ float segment_mid(const MidIndex<ND, DIAG>& mi, const float* src,
                                                      const float* vw, float* p0, float t1, float* c) {
    float p1[ND];
    float ss = 0.0f;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        p1[a] = src[a] + t1 * vw[a];
        ss = add_sq(ss, p1[a] - p0[a], a == 0);
    }
    const float len = seg_sqrt(ss);
    mi.from_ends(p0, p1, c);
#pragma unroll
    for (int a = 0; a < ND; ++a) p0[a] = p1[a];
    return len;
}

template <int ND, bool DIAG>
__host__ __device__ __forceinline__ bool segment_from_t(const Grid<ND>& g, const MidIndex<ND, DIAG>& mi,
                                                        const float* src, const float* vw, float* p0, float t1,
                                                        uint32_t* idx, float* len) {
    float c[ND];
    *len = segment_mid<ND, DIAG>(mi, src, vw, p0, t1, c);
    return flat_voxel<ND>(g, c, idx);
}

}  // namespace ctlab
