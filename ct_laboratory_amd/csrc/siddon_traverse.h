// siddon_traverse.h -- one-ray-per-lane Siddon traversal for gfx950.
//
// The reference defines a ray's segments by enumerating every voxel-plane
// crossing, sorting the list and walking consecutive pairs
// (ct_laboratory/tomography/ct_projector_3d_torch.py:78-158 and :161-274;
// src/ct_projector_3d.cu:834-1012).  Here the per-axis crossing sequences are
// each monotone in the plane index, so their merge *is* the sorted list: the
// lane keeps one pending crossing per axis and pops the smallest, with no
// scratch array and no sort.
//
// Bit-exactness.  Every crossing is evaluated with the reference's formula
// (torch: RN(RN(p-0.5-s)/dir); otf: RN(RN(p-0.5-s)*RN(1/dir))) and every
// segment's voxel with the reference's midpoint rule and operation order, so
// the (voxel, length) sequence equals the reference's for each segment.  The
// file is compiled with floating-point contraction off; fused multiply-adds
// appear only where written as __builtin_fmaf (the otf forms, which follow
// nvcc's default contraction of the CUDA source).
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

#pragma clang fp contract(off)

namespace ctlab {

enum : int { MODE_TORCH = 0, MODE_OTF = 1 };
enum : int { ZERO_THRESH = 0, ZERO_EXACT = 1 };

// Uniform per-launch geometry.  minv is row-major ND x ND.
template <int ND>
struct Grid {
    int n[ND];
    float minv[ND * ND];
    float b[ND];
    int diag;       // off-diagonal entries of minv are all zero
    int zero_rule;  // ZERO_THRESH: |dir| < 1e-12 has no crossings; ZERO_EXACT: dir == 0
};

template <int ND>
__host__ __device__ __forceinline__ Grid<ND> load_grid(const int* n, const float* __restrict__ minv,
                                              const float* __restrict__ b, int zero_rule) {
    Grid<ND> g;
    int diag = 1;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        g.n[a] = n[a];
        g.b[a] = b[a];
    }
#pragma unroll
    for (int i = 0; i < ND * ND; ++i) {
        g.minv[i] = minv[i];
        if ((i / ND) != (i % ND) && g.minv[i] != 0.0f) diag = 0;
    }
    g.diag = diag;
    g.zero_rule = zero_rule;
    return g;
}

// Reference voxel-space transform of an endpoint: (p - b) @ M_inv^T with
// separate rounding (3d_torch.py:100-103), or nvcc's contraction of
// apply_affine_inverse_3d (3d.cu:21-36) in otf mode.
template <int ND, int MODE>
__host__ __device__ __forceinline__ void to_voxel(const Grid<ND>& g, const float* p, float* out) {
    float sh[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) sh[j] = p[j] - g.b[j];
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        if (MODE == MODE_OTF) {
            float acc = g.minv[a * ND] * sh[0];
#pragma unroll
            for (int j = 1; j < ND; ++j) acc = __builtin_fmaf(g.minv[a * ND + j], sh[j], acc);
            out[a] = acc;
        } else {
            float acc = sh[0] * g.minv[a * ND];
#pragma unroll
            for (int j = 1; j < ND; ++j) acc = acc + sh[j] * g.minv[a * ND + j];
            out[a] = acc;
        }
    }
}

// Midpoint voxel coordinates (3d_torch.py:244-250; 3d.cu:985-991).
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

// Round half to even and range check; returns the flat index or -1.
template <int ND>
__host__ __device__ __forceinline__ int flat_voxel(const Grid<ND>& g, const float* c) {
    int idx = 0;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        float r = __builtin_rintf(c[a]);  // v_rndne_f32: ties to even
        ok = ok && (r >= 0.0f) && (r < (float)g.n[a]);
        idx = idx * g.n[a] + (int)r;
    }
    return ok ? idx : -1;
}

// Plane crossing of axis a at plane index p.
template <int MODE>
__host__ __device__ __forceinline__ float crossing(int p, float s, float dir_or_inv) {
    float num = ((float)p - 0.5f) - s;
    if (MODE == MODE_OTF) return num * dir_or_inv;
    return num / dir_or_inv;
}

// Visits every segment of one ray whose rounded midpoint lies in the grid,
// in increasing t, as visit(flat_index, length).  Zero-length segments are
// not visited (they contribute exactly zero to both projections).
template <int ND, int MODE, bool DIAG, class Visit>
__host__ __device__ __forceinline__ void trace_ray(const Grid<ND>& g, const float* src, const float* dst,
                                          Visit&& visit) {
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

    to_voxel<ND, MODE>(g, src, sv);
    to_voxel<ND, MODE>(g, dst, dv);

    float dirq[ND];  // divisor (torch) or reciprocal (otf) per axis
    float next[ND];
    int pl[ND], step[ND];
    bool act[ND];
    float t_lo = 0.0f, t_hi = 1.0f;
    const float kInf = __builtin_huge_valf();

#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float dir = dv[a] - sv[a];
        // Rounding-error bound of a midpoint coordinate on this axis, voxel units.
        float mag = fabsf(sv[a]) + fabsf(dv[a]) + (float)g.n[a] + 1.0f;
#pragma unroll
        for (int j = 0; j < ND; ++j)
            mag += fabsf(g.minv[a * ND + j]) * (fabsf(src[j]) + fabsf(dst[j]) + fabsf(g.b[j]));
        const float E = mag * 7.62939453125e-06f;  // 2^-17
        bool active;
        if (MODE == MODE_OTF) {
            active = fabsf(dir) > 1.0e-12f;
        } else {
            active = (g.zero_rule == ZERO_EXACT) ? (dir != 0.0f) : !(fabsf(dir) < 1.0e-12f);
        }
        act[a] = active;
        if (!active) {
            // Constant coordinate: a ray outside the slab by more than E has no
            // in-range segment at all.
            if (sv[a] < -0.5f - E || sv[a] > (float)g.n[a] - 0.5f + E) return;
            dirq[a] = 1.0f;
            step[a] = 0;
            continue;
        }
        dirq[a] = (MODE == MODE_OTF) ? 1.0f / dir : dir;
        const bool up = dir > 0.0f;
        step[a] = up ? 1 : -1;
        const float t_in = crossing<MODE>(up ? 0 : g.n[a], sv[a], dirq[a]);
        const float t_out = crossing<MODE>(up ? g.n[a] : 0, sv[a], dirq[a]);
        const float tau = 4.0f * E / fabsf(dir) + 9.5367431640625e-07f;  // + 2^-20
        t_lo = fmaxf(t_lo, t_in - tau);
        t_hi = fminf(t_hi, t_out + tau);
    }
    if (!(t_lo <= t_hi)) return;

    // Jump each axis to its first crossing >= t_lo (crossings are monotone).
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        next[a] = kInf;
        pl[a] = 0;
        if (!act[a]) continue;
        const int n = g.n[a];
        const float c = sv[a] + t_lo * (dv[a] - sv[a]);
        int p;
        // The clamps are written so that a NaN guess lands inside [-1, n+1]:
        // both search loops are then bounded by the grid size.
        if (step[a] > 0) {
            float f = ceilf(c + 0.5f);
            p = (f >= 0.0f) ? ((f <= (float)(n + 1)) ? (int)f : n + 1) : 0;
            while (p > 0 && crossing<MODE>(p - 1, sv[a], dirq[a]) >= t_lo) --p;
            while (p <= n && crossing<MODE>(p, sv[a], dirq[a]) < t_lo) ++p;
        } else {
            float f = floorf(c + 0.5f);
            p = (f >= -1.0f) ? ((f <= (float)n) ? (int)f : n) : -1;
            while (p < n && crossing<MODE>(p + 1, sv[a], dirq[a]) >= t_lo) ++p;
            while (p >= 0 && crossing<MODE>(p, sv[a], dirq[a]) < t_lo) --p;
        }
        pl[a] = p;
        if (p >= 0 && p <= n) {
            float t = crossing<MODE>(p, sv[a], dirq[a]);
            next[a] = (t > 1.0f) ? kInf : t;
        }
    }

    auto pop = [&](float& t) -> bool {
        float m = next[0];
#pragma unroll
        for (int a = 1; a < ND; ++a) m = fminf(m, next[a]);
        if (!(m <= t_hi)) return false;
        int ax = ND - 1;
#pragma unroll
        for (int a = ND - 2; a >= 0; --a)
            if (next[a] == m) ax = a;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            if (a == ax) {
                int p = pl[a] + step[a];
                pl[a] = p;
                float tn = kInf;
                if (p >= 0 && p <= g.n[a]) {
                    tn = crossing<MODE>(p, sv[a], dirq[a]);
                    if (tn > 1.0f) tn = kInf;
                }
                next[a] = tn;
            }
        }
        t = m;
        return true;
    };

    float t0;
    if (MODE == MODE_OTF && t_lo == 0.0f) {
        t0 = 0.0f;  // endpoint inserted by collect_hits (3d.cu:890)
    } else if (!pop(t0)) {
        return;
    }
    const bool close_at_one = (MODE == MODE_OTF) && (t_hi == 1.0f);

    float p0[ND];
#pragma unroll
    for (int a = 0; a < ND; ++a) p0[a] = src[a] + t0 * vw[a];

    while (true) {
        float t1;
        bool last = false;
        if (!pop(t1)) {
            if (close_at_one && t0 < 1.0f) {
                t1 = 1.0f;  // endpoint inserted by collect_hits (3d.cu:891)
                last = true;
            } else {
                break;
            }
        }
        float c[ND];
        float len;
        if (MODE == MODE_TORCH) {
            float p1[ND], m[ND];
            float ss = 0.0f;
#pragma unroll
            for (int a = 0; a < ND; ++a) {
                p1[a] = src[a] + t1 * vw[a];
                const float dd = p1[a] - p0[a];
                const float sq = dd * dd;
                ss = (a == 0) ? sq : ss + sq;
                m[a] = 0.5f * (p0[a] + p1[a]);
                p0[a] = p1[a];
            }
            len = __builtin_sqrtf(ss);
            mid_coords<ND, MODE, DIAG>(g, m, c);
        } else {
            float m[ND];
            const float tm = 0.5f * (t0 + t1);
#pragma unroll
            for (int a = 0; a < ND; ++a) m[a] = __builtin_fmaf(tm, vw[a], src[a]);
            len = (t1 > t0) ? (t1 - t0) * rl : 0.0f;
            mid_coords<ND, MODE, DIAG>(g, m, c);
        }
        t0 = t1;
        if (len > 0.0f) {
            const int idx = flat_voxel<ND>(g, c);
            if (idx >= 0) visit(idx, len);
        }
        if (last) break;
    }
}

// Emits, in increasing order, every plane crossing t in [0, 1] of the whole
// ray: the sorted list of compute_intersections_{2d,3d}
// (ct_projector_3d_torch.py:78-158, src/ct_projector_3d.cu:42-96) without its
// INF padding.  Crossings use the division form RN(RN(p-0.5-s)/dir); an axis
// has none when the zero rule says so.  Returns the number emitted.
template <int ND, class Emit>
__host__ __device__ __forceinline__ int emit_crossings(const Grid<ND>& g, const float* src, const float* dst,
                                                       Emit&& emit) {
    float sv[ND], dv[ND], dir[ND], next[ND];
    int pl[ND], step[ND];
    const float kInf = __builtin_huge_valf();
    to_voxel<ND, MODE_TORCH>(g, src, sv);
    to_voxel<ND, MODE_TORCH>(g, dst, dv);
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        dir[a] = dv[a] - sv[a];
        next[a] = kInf;
        pl[a] = 0;
        step[a] = 0;
        const bool active = (g.zero_rule == ZERO_EXACT) ? (dir[a] != 0.0f) : !(fabsf(dir[a]) < 1.0e-12f);
        if (!active) continue;
        const int n = g.n[a];
        int p;
        // First crossing with t >= 0 in traversal order (monotone in p).
        if (dir[a] > 0.0f) {
            step[a] = 1;
            const float f = ceilf(sv[a] + 0.5f);
            p = (f >= 0.0f) ? ((f <= (float)(n + 1)) ? (int)f : n + 1) : 0;
            while (p > 0 && crossing<MODE_TORCH>(p - 1, sv[a], dir[a]) >= 0.0f) --p;
            while (p <= n && crossing<MODE_TORCH>(p, sv[a], dir[a]) < 0.0f) ++p;
        } else {
            step[a] = -1;
            const float f = floorf(sv[a] + 0.5f);
            p = (f >= -1.0f) ? ((f <= (float)n) ? (int)f : n) : -1;
            while (p < n && crossing<MODE_TORCH>(p + 1, sv[a], dir[a]) >= 0.0f) ++p;
            while (p >= 0 && crossing<MODE_TORCH>(p, sv[a], dir[a]) < 0.0f) --p;
        }
        pl[a] = p;
        if (p >= 0 && p <= n) {
            const float t = crossing<MODE_TORCH>(p, sv[a], dir[a]);
            next[a] = (t > 1.0f) ? kInf : t;
        }
    }
    int count = 0;
    while (true) {
        float m = next[0];
#pragma unroll
        for (int a = 1; a < ND; ++a) m = fminf(m, next[a]);
        if (!(m <= 1.0f)) break;
        int ax = ND - 1;
#pragma unroll
        for (int a = ND - 2; a >= 0; --a)
            if (next[a] == m) ax = a;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            if (a == ax) {
                const int p = pl[a] + step[a];
                pl[a] = p;
                float tn = kInf;
                if (p >= 0 && p <= g.n[a]) {
                    tn = crossing<MODE_TORCH>(p, sv[a], dir[a]);
                    if (tn > 1.0f) tn = kInf;
                }
                next[a] = tn;
            }
        }
        emit(m);
        ++count;
    }
    return count;
}

// Segment geometry of one pair of crossings given as t values, in the op order
// of the torch consumer (3d_torch.py:219-271): world points src + t (dst-src),
// length |P1 - P0|, midpoint voxel by round-half-even.  p0 holds P(t0) on
// entry and P(t1) on return.  Returns the flat voxel or -1; *len receives the
// length.
template <int ND, bool DIAG>
__host__ __device__ __forceinline__ int segment_from_t(const Grid<ND>& g, const float* src, const float* vw,
                                                       float* p0, float t1, float* len) {
    float p1[ND], m[ND], c[ND];
    float ss = 0.0f;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        p1[a] = src[a] + t1 * vw[a];
        const float dd = p1[a] - p0[a];
        const float sq = dd * dd;
        ss = (a == 0) ? sq : ss + sq;
        m[a] = 0.5f * (p0[a] + p1[a]);
        p0[a] = p1[a];
    }
    *len = __builtin_sqrtf(ss);
    mid_coords<ND, MODE_TORCH, DIAG>(g, m, c);
    return flat_voxel<ND>(g, c);
}

}  // namespace ctlab
