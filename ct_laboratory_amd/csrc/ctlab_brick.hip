// ctlab_brick.hip -- atomic-free 3D back projection for rays that come as a
// stack of flat-panel views ([view][rows][cols], one source per view), the
// layout of the cone-beam and static-ring builders.
//
// A workgroup owns a 32x32x8 brick of the output volume, accumulated in LDS
// (64 KB, two workgroups per CU), and loops over every view.  Per view it projects the brick's
// corners from the view's source onto the detector plane and takes the padded
// bounding rectangle of detector pixels (the brick's shadow); every ray whose
// rounded voxels touch the brick lies in that rectangle.  Each such ray is
// traced with the exact traversal clipped to the brick window
// (trace_ray_window): the brick receives exactly the segments whose voxel it
// contains.  The brick is written once with plain stores: no global atomics.
//
// Accumulation is 64-bit fixed point (ds_add_u64), because LDS float atomic
// adds run at a small fraction of integer atomic throughput on gfx950.  Each
// brick first takes the largest |sinogram| value over its rays; with the
// longest possible segment (the voxel's body diagonal) and the number of rays
// this bounds every voxel's sum, and a power-of-two scale maps that bound just
// below 2^62.  Each float contribution sino*len (the reference's product) is
// scaled exactly and rounded to an integer: the quantum is 2^-62 of the bound,
// far below float accumulation error, and integer addition makes the result
// independent of the order of the rays (deterministic).  Bricks whose rays
// carry non-finite values fall back to float atomics into global memory so
// that INF/NaN propagate as in the reference.
//
// Exactness does not depend on the detector model: the frame of a view is only
// used to bound the shadow, conservatively.  A view whose rays do not fit the
// flat-panel model within a quarter pixel (source moving, detector not
// affine) is marked invalid and uses its whole detector for every brick.
#include <mutex>

#include "ctlab_common.h"
#include "fixed_point.h"
#include "siddon_traverse.h"

namespace ctlab {

constexpr int kBrickX = 32, kBrickY = 32, kBrickZ = 8;  // brick edges, voxels
constexpr int kBrickThreads = 1024;                      // 16 waves
constexpr int kFrameFloats = 16;
// LDS strides of the brick (64-bit cells); the x stride 32*8+1 is odd so that
// lanes stepping along x hit different banks.  About 64 KB, so two workgroups
// (32 waves, 8 per SIMD) with their ray queues share a CU's 160 KB of LDS.
// Rays of a circular scan run mostly in the x-y plane: a brick long in x and y
// and thin in z is entered by fewer rays per segment than a cubic one.
constexpr int kStrideZ = 1, kStrideY = kBrickZ, kStrideX = kBrickY * kBrickZ + 1;
constexpr int kBrickLds = kBrickX * kStrideX;

struct Dims3 {
    int n[3];
};

// Frame layout: S[3] source, O[3] pixel (0,0), R[3] row step, C[3] column step
// (all voxel coordinates), E bound, source deviation, pixel deviation, valid.
enum : int { F_S = 0, F_O = 3, F_R = 6, F_C = 9, F_E = 12, F_DEVS = 13, F_DEVD = 14, F_VALID = 15 };

__device__ __forceinline__ void to_voxel_f(const Grid<3>& g, const float* p, float* out) {
    to_voxel<3, MODE_TORCH>(g, p, out);
}

// Block-wide max of a float (1024 threads max).
__device__ float block_max(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) red[w] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) / 64;
    if (threadIdx.x < 64) {
        float x = (l < nw) ? red[l] : -__builtin_huge_valf();
        for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
        if (l == 0) red[0] = x;
    }
    __syncthreads();
    const float r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_cone_frames(const float* __restrict__ src, const float* __restrict__ dst,
                                                     int rows, int cols, Dims3 dims,
                                                     const float* __restrict__ minv, const float* __restrict__ b,
                                                     float* __restrict__ frames) {
    __shared__ float red[16];
    const int64_t v = blockIdx.x;
    const int64_t base = v * rows * cols;
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    float S[3], O[3], R[3], C[3], t0[3], t1[3];
    to_voxel_f(g, src + 3 * base, S);
    to_voxel_f(g, dst + 3 * base, O);
    to_voxel_f(g, dst + 3 * (base + (int64_t)(cols - 1)), t0);
    to_voxel_f(g, dst + 3 * (base + (int64_t)(rows - 1) * cols), t1);
    for (int a = 0; a < 3; ++a) {
        C[a] = cols > 1 ? (t0[a] - O[a]) / (float)(cols - 1) : 0.0f;
        R[a] = rows > 1 ? (t1[a] - O[a]) / (float)(rows - 1) : 0.0f;
    }
    float devs = 0.0f, devd = 0.0f, emax = 0.0f;
    const int64_t npix = (int64_t)rows * cols;
    for (int64_t i = threadIdx.x; i < npix; i += blockDim.x) {
        const int r = (int)(i / cols), c = (int)(i % cols);
        float sv[3], dv[3];
        to_voxel_f(g, src + 3 * (base + i), sv);
        to_voxel_f(g, dst + 3 * (base + i), dv);
        for (int a = 0; a < 3; ++a) {
            devs = fmaxf(devs, fabsf(sv[a] - S[a]));
            devd = fmaxf(devd, fabsf(dv[a] - (O[a] + (float)r * R[a] + (float)c * C[a])));
            // the traversal's rounding-error bound (siddon_traverse.h)
            float mag = fabsf(sv[a]) + fabsf(dv[a]) + (float)g.n[a] + 1.0f;
            for (int j = 0; j < 3; ++j)
                mag += fabsf(g.minv[a * 3 + j]) *
                       (fabsf(src[3 * (base + i) + j]) + fabsf(dst[3 * (base + i) + j]) + fabsf(g.b[j]));
            emax = fmaxf(emax, mag * 7.62939453125e-06f);
        }
    }
    devs = block_max(devs, red);
    devd = block_max(devd, red);
    emax = block_max(emax, red);
    if (threadIdx.x == 0) {
        float* f = frames + v * kFrameFloats;
        for (int a = 0; a < 3; ++a) {
            f[F_S + a] = S[a];
            f[F_O + a] = O[a];
            f[F_R + a] = R[a];
            f[F_C + a] = C[a];
        }
        const float pr = sqrtf(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        const float pc = sqrtf(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
        const float pitch = fminf(pr, pc);
        f[F_E] = emax;
        f[F_DEVS] = devs;
        f[F_DEVD] = devd;
        const bool ok = rows > 1 && cols > 1 && pitch > 0.0f && devs <= 0.25f * pitch && devd <= 0.25f * pitch &&
                        emax <= 1.0f;
        f[F_VALID] = ok ? 1.0f : 0.0f;
    }
}

// Shadow rectangle of the box [lo-0.5-m, lo+ext-0.5+m] on one view's detector.
// Returns false when the frame cannot bound it (then the caller uses the whole
// detector).
__device__ bool shadow_rect(const float* __restrict__ f, const int* lo, const int* ext, int rows, int cols, int* r0,
                            int* r1, int* c0, int* c1) {
    if (f[F_VALID] == 0.0f) return false;
    // A segment whose rounded voxel is in the box has its computed midpoint in
    // the box's cells [lo-0.5, lo+ext-0.5], so the ray passes within E (the
    // midpoint error bound) of them; between source and detector the model
    // line S -> O + rR + cC of its pixel deviates from the ray by at most the
    // larger of the source and pixel deviations.  The rays are the pixel
    // centres, so the pixels needed are the integer points of the projected
    // box.  Slack for float rounding: 0.05 voxel on the box (the voxel
    // coordinates of the frame and corners carry errors of ~1e-4) and a
    // quarter pixel on the rectangle (projection errors of ~1e-4 pixel).
    const float m = f[F_E] + f[F_DEVS] + f[F_DEVD] + 0.05f;
    const float S[3] = {f[F_S], f[F_S + 1], f[F_S + 2]};
    const float B[3] = {-f[F_R], -f[F_R + 1], -f[F_R + 2]};
    const float Cm[3] = {-f[F_C], -f[F_C + 1], -f[F_C + 2]};
    const float rhs[3] = {f[F_O] - S[0], f[F_O + 1] - S[1], f[F_O + 2] - S[2]};
    auto cross = [](const float* a, const float* c, float* o) {
        o[0] = a[1] * c[2] - a[2] * c[1];
        o[1] = a[2] * c[0] - a[0] * c[2];
        o[2] = a[0] * c[1] - a[1] * c[0];
    };
    auto dot = [](const float* a, const float* c) { return a[0] * c[0] + a[1] * c[1] + a[2] * c[2]; };
    float bxc[3], rxc[3], bxr[3];
    cross(B, Cm, bxc);
    cross(rhs, Cm, rxc);
    cross(B, rhs, bxr);
    const float num_l = dot(rhs, bxc);
    float rmin = 3e38f, rmax = -3e38f, cmin = 3e38f, cmax = -3e38f;
    for (int k = 0; k < 8; ++k) {
        float A[3];
        for (int a = 0; a < 3; ++a) {
            const float x = ((k >> a) & 1) ? (float)(lo[a] + ext[a]) - 0.5f + m : (float)lo[a] - 0.5f - m;
            A[a] = x - S[a];
        }
        const float det = dot(A, bxc);
        if (!(fabsf(det) > 1e-20f)) return false;
        if (!(num_l / det > 0.0f)) return false;  // corner not in front of the source
        const float inv = 1.0f / det;
        const float r = dot(A, rxc) * inv;
        const float c = dot(A, bxr) * inv;
        rmin = fminf(rmin, r);
        rmax = fmaxf(rmax, r);
        cmin = fminf(cmin, c);
        cmax = fmaxf(cmax, c);
    }
    const float pad = 0.25f;
    const float fr0 = floorf(rmin - pad), fr1 = ceilf(rmax + pad), fc0 = floorf(cmin - pad), fc1 = ceilf(cmax + pad);
    if (!(fr1 >= 0.0f) || !(fc1 >= 0.0f) || !(fr0 <= (float)(rows - 1)) || !(fc0 <= (float)(cols - 1))) {
        *r0 = 0;
        *r1 = -1;  // empty
        *c0 = 0;
        *c1 = -1;
        return true;
    }
    *r0 = (int)fmaxf(fr0, 0.0f);
    *r1 = (int)fminf(fr1, (float)(rows - 1));
    *c0 = (int)fmaxf(fc0, 0.0f);
    *c1 = (int)fminf(fc1, (float)(cols - 1));
    return true;
}

// Brick plan: for every brick and view, the shadow rectangle (r0, c0, width)
// and the exclusive prefix of the rectangle sizes over the views, so that the
// back projector maps a flat item number to (view, pixel); then the total per
// brick, the midpoint error bound E of every view (float) and a copy of the
// views' frames.  Depends only on the geometry: computed once, reused by every
// projection.
constexpr int kPlanInts = 4;  // r0, c0, w, prefix

struct BrickGrid {
    int nbx, nby, nbz;
};

__device__ __forceinline__ Window<3> brick_window(const Dims3& dims, const BrickGrid& bg, int bid) {
    const int bz = bid % bg.nbz, by = (bid / bg.nbz) % bg.nby, bx = bid / (bg.nbz * bg.nby);
    Window<3> win;
    win.lo[0] = bx * kBrickX;
    win.lo[1] = by * kBrickY;
    win.lo[2] = bz * kBrickZ;
    win.ext[0] = min(kBrickX, dims.n[0] - win.lo[0]);
    win.ext[1] = min(kBrickY, dims.n[1] - win.lo[1]);
    win.ext[2] = min(kBrickZ, dims.n[2] - win.lo[2]);
    win.stride[0] = kStrideX;
    win.stride[1] = kStrideY;
    win.stride[2] = kStrideZ;
    return win;
}

__global__ __launch_bounds__(256) void k_brick_plan(Dims3 dims, BrickGrid bg, int n_view, int rows, int cols,
                                                    const float* __restrict__ frames, int32_t* __restrict__ plan,
                                                    uint32_t* __restrict__ totals, float* __restrict__ view_e,
                                                    float* __restrict__ frames_copy) {
    __shared__ uint32_t s_wsum[4];
    const int bid = blockIdx.x;
    if (bid == 0) {
        for (int v = threadIdx.x; v < n_view; v += blockDim.x) view_e[v] = frames[(int64_t)v * kFrameFloats + F_E];
        for (int64_t i = threadIdx.x; i < (int64_t)n_view * kFrameFloats; i += blockDim.x) frames_copy[i] = frames[i];
    }
    const Window<3> win = brick_window(dims, bg, bid);
    int32_t* __restrict__ p = plan + (int64_t)bid * n_view * kPlanInts;
    const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    uint32_t running = 0;
    for (int v0 = 0; v0 < n_view; v0 += 256) {
        const int v = v0 + tid;
        int r0 = 0, r1 = -1, c0 = 0, c1 = -1;
        if (v < n_view) {
            if (!shadow_rect(frames + (int64_t)v * kFrameFloats, win.lo, win.ext, rows, cols, &r0, &r1, &c0, &c1)) {
                r0 = 0;
                r1 = rows - 1;
                c0 = 0;
                c1 = cols - 1;
            }
        }
        const int h = r1 - r0 + 1, w = c1 - c0 + 1;
        const uint32_t cnt = (v < n_view && h > 0 && w > 0) ? (uint32_t)h * (uint32_t)w : 0u;
        uint32_t x = cnt;  // inclusive scan over the block
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_wsum[wv] = x;
        __syncthreads();
        uint32_t base = running;
        for (int k = 0; k < wv; ++k) base += s_wsum[k];
        if (v < n_view) {
            p[(int64_t)v * kPlanInts + 0] = r0;
            p[(int64_t)v * kPlanInts + 1] = c0;
            p[(int64_t)v * kPlanInts + 2] = cnt ? w : 0;
            p[(int64_t)v * kPlanInts + 3] = (int32_t)(base + x - cnt);
        }
        running += s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
        __syncthreads();
    }
    if (tid == 0) totals[bid] = running;
}

// Upper estimate of the number of plane crossings of a ray inside the box
// [blo, bhi] (voxel coordinates, the brick's cells widened by at least the
// midpoint error bound), 0 when the ray misses the box on t in [0, 1].  Any
// ray with a segment whose rounded voxel is in the brick passes through the
// widened box (its computed midpoint is within E of the exact line), so 0 is
// returned only for rays the exact traversal would not visit; the extra
// 2^-14 relative slack covers the rounding of this estimate itself.
__device__ __forceinline__ uint32_t slab_estimate(const float* sv, const float* dv, const float* blo,
                                                  const float* bhi, float extra) {
    float tin = 0.0f, tout = 1.0f, span = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float dir = dv[a] - sv[a];
        const float slack = extra + 6.103515625e-05f * (fabsf(sv[a]) + fabsf(dv[a]));
        const float lo = blo[a] - slack, hi = bhi[a] + slack;
        if (dir == 0.0f) {
            if (!(sv[a] >= lo && sv[a] <= hi)) return 0u;
            continue;
        }
        const float rc = __builtin_amdgcn_rcpf(dir);
        const float t1 = (lo - sv[a]) * rc, t2 = (hi - sv[a]) * rc;
        tin = fmaxf(tin, fminf(t1, t2));
        tout = fminf(tout, fmaxf(t1, t2));
        span += fabsf(dir);
    }
    if (!(tin <= tout)) return 0u;
    return (uint32_t)fminf(span * (tout - tin), 1.0e6f) + 2u;
}

// From the ray's own end points.
template <int MODE>
__device__ __forceinline__ uint32_t segment_estimate(const Grid<3>& g, const float* __restrict__ s,
                                                     const float* __restrict__ d, const float* blo,
                                                     const float* bhi) {
    float sv[3], dv[3];
    to_voxel<3, MODE>(g, s, sv);
    to_voxel<3, MODE>(g, d, dv);
    return slab_estimate(sv, dv, blo, bhi, 0.0f);
}

// From the view's flat-panel frame: the model line from the view's source to
// O + r R + c C deviates from the ray by at most the source and pixel
// deviations, which widen the box (the same float expression as the frame's
// fit, so the bound holds for these very values).
__device__ __forceinline__ uint32_t frame_estimate(const float* __restrict__ f, int r, int c, const float* blo,
                                                   const float* bhi) {
    const float sv[3] = {f[F_S], f[F_S + 1], f[F_S + 2]};
    float dv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dv[a] = f[F_O + a] + (float)r * f[F_R + a] + (float)c * f[F_C + a];
    return slab_estimate(sv, dv, blo, bhi, f[F_DEVS] + f[F_DEVD]);
}

// Items are traced in chunks of at most kChunk.  Per chunk the plan entries
// and frames of the next kStage views are staged in LDS (one load latency for
// the block instead of a dependent chain per item); every thread then takes
// two items, finds their views by a binary search in LDS, drops rays with a
// zero value or no crossing of the brick, and files the rest into buckets by
// their estimated crossing count (quarter octaves); then the chunk's rays are
// traced in bucket order, so the 64 lanes of a wave trace rays of similar
// length instead of idling behind the longest one.  A chunk ends at the first
// item beyond the staged views.
constexpr int kPerThread = 4;  // items per thread and chunk
constexpr int kChunk = kPerThread * kBrickThreads;  // queue entries are chunk offsets (< 2^12)
constexpr int kBuckets = 32;  // quarter-octave length classes (1, 1.25, 1.5, 1.75, 2, 2.5, ... crossings)
constexpr int kStage = 64;

// local / w and local % w of a decoded item.  For local < 2^22 a float
// reciprocal gives a quotient estimate within 1/2 of local / w (relative error
// below 2^-22; exact for w = 1), so its truncation is off by at most one, fixed
// by one correction of the remainder; larger values use integer division.
__device__ __forceinline__ uint32_t div_rect(uint32_t a, uint32_t b, uint32_t* rem) {
    if (a < (1u << 22)) {
        uint32_t q = (uint32_t)((float)a * __builtin_amdgcn_rcpf((float)b));
        int32_t r = (int32_t)(a - q * b);
        if (r < 0) {
            --q;
            r += (int32_t)b;
        } else if (r >= (int32_t)b) {
            ++q;
            r -= (int32_t)b;
        }
        *rem = (uint32_t)r;
        return q;
    }
    *rem = a % b;
    return a / b;
}

// Largest j in [0, n) with prefix(j) <= item, for prefixes non-decreasing in j
// and prefix(0) <= item.
__device__ __forceinline__ int staged_view(const int4* s_plan, int n, uint32_t item) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((uint32_t)s_plan[mid].w <= item) lo = mid; else hi = mid;
    }
    return lo;
}

// Traces every ray of the brick and hands (window index, contribution) to
// add().  Each sinogram value is first multiplied by s1*s2 (powers of two, so
// exactly, and the product val*len scales with it).  e_brick bounds the
// midpoint error of every ray of the brick (non-finite or above 8: unknown).
// s_plan holds kStage + 1 plan entries and s_frame kStage frames (LDS).
template <int MODE, bool DIAG, class Add>
__device__ __forceinline__ void brick_items(const Grid<3>& g, const Window<3>& win, const float* __restrict__ sn,
                                            const float* __restrict__ src, const float* __restrict__ dst,
                                            const int32_t* __restrict__ p, const float* __restrict__ frames,
                                            int n_view, int rows, int cols, uint32_t item_lo, uint32_t item_hi,
                                            uint32_t total, float e_brick, float s1,
                                            float s2, uint16_t* queue, uint32_t* bcnt, int4* s_plan, float* s_frame,
                                            Add&& add) {
    const int tid = threadIdx.x;
    const bool e_ok = e_brick <= 8.0f;
    const float e_hint = e_ok ? e_brick : -1.0f;
    const float m = (e_ok ? e_brick : 8.0f) + 0.05f;
    float blo[3], bhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        blo[a] = (float)win.lo[a] - 0.5f - m;
        bhi[a] = (float)(win.lo[a] + win.ext[a]) - 0.5f + m;
    }
    // view of item_lo: the largest view whose prefix is <= item_lo (block-uniform)
    int vs = 0;
    {
        int lo = 0, hi = n_view;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((uint32_t)p[mid * kPlanInts + 3] <= item_lo) lo = mid; else hi = mid;
        }
        vs = lo;
    }
    uint32_t c0 = item_lo;
    while (c0 < item_hi) {
        if (tid <= kBuckets) bcnt[tid] = 0u;
        if (tid <= kStage) {
            const int v = vs + tid;
            int4 e = make_int4(0, 0, 1, (int)total);  // past the last view: prefix = total
            if (v < n_view) e = *reinterpret_cast<const int4*>(p + (int64_t)v * kPlanInts);
            s_plan[tid] = e;
            if (tid < kStage && v < n_view) {
                const float* __restrict__ f = frames + (int64_t)v * kFrameFloats;
#pragma unroll
                for (int k = 0; k < kFrameFloats; ++k) s_frame[tid * kFrameFloats + k] = f[k];
            }
        }
        __syncthreads();
        const uint32_t lim = (uint32_t)s_plan[kStage].w;  // first item beyond the staged views
        const uint32_t c1 = min(c0 + (uint32_t)kChunk, min(lim, item_hi));
        // item -> (staged view j, detector row r, column c, ray)
        auto decode = [&](uint32_t item, int& j, int& r, int& c) -> uint32_t {
            j = staged_view(s_plan, kStage, item);
            const int4 e = s_plan[j];
            const uint32_t local = item - (uint32_t)e.w, w = (uint32_t)e.z;
            uint32_t rem;
            r = e.x + (int)div_rect(local, w, &rem);
            c = e.y + (int)rem;
            return (uint32_t)(((int64_t)(vs + j) * rows + r) * cols + c);
        };
        uint32_t key[kPerThread];
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            key[k] = ~0u;
            const uint32_t item = c0 + (uint32_t)tid + (uint32_t)(k * kBrickThreads);
            if (item < c1) {
                int j, r, c;
                const uint32_t ray = decode(item, j, r, c);
                if (sn[ray] != 0.0f) {
                    const float* f = s_frame + j * kFrameFloats;
                    const uint32_t est = (f[F_VALID] != 0.0f)
                                             ? frame_estimate(f, r, c, blo, bhi)
                                             : segment_estimate<MODE>(g, src + 3ull * ray, dst + 3ull * ray, blo, bhi);
                    if (est) {
                        const uint32_t l2 = 31u - (uint32_t)__clz(est);  // est in [2^l2, 2^(l2+1))
                        const uint32_t bk = min(4u * l2 + ((est << 2) >> l2) - 4u, (uint32_t)(kBuckets - 1));
                        key[k] = (bk << 12) | atomicAdd(&bcnt[bk], 1u);
                    }
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t run = 0;
            for (int k = 0; k < kBuckets; ++k) {
                const uint32_t c = bcnt[k];
                bcnt[k] = run;
                run += c;
            }
            bcnt[kBuckets] = run;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPerThread; ++k)
            if (key[k] != ~0u)
                queue[bcnt[key[k] >> 12] + (key[k] & 0xfffu)] = (uint16_t)(tid + k * kBrickThreads);
        __syncthreads();
        const uint32_t hits = bcnt[kBuckets];
        for (uint32_t q = tid; q < hits; q += kBrickThreads) {
            int j, r, c;
            const uint32_t ray = decode(c0 + queue[q], j, r, c);
            const float val = sn[ray] * s1 * s2;
            float s[3], d[3];
            for (int a = 0; a < 3; ++a) {
                s[a] = src[3ull * ray + a];
                d[a] = dst[3ull * ray + a];
            }
            trace_ray_window<3, MODE, DIAG, false>(
                g, win, s, d, [&](uint32_t li, float len) { add(li, val, len); }, e_hint);
        }
        // the next chunk starts in the largest staged view whose prefix <= c1
        vs += staged_view(s_plan, kStage + 1, c1);
        c0 = c1;
        __syncthreads();
    }
}

// Longest segment any ray can have inside one voxel: the largest body
// diagonal of the voxel parallelepiped, |M d| for d in {(1,1,1), (1,1,-1),
// (1,-1,1), (-1,1,1)}, with M the inverse of M_inv.  Evaluated in double.
__device__ __forceinline__ double voxel_diagonal(const Grid<3>& g) {
    const float* a = g.minv;
    const double c00 = (double)a[4] * a[8] - (double)a[5] * a[7], c01 = (double)a[5] * a[6] - (double)a[3] * a[8],
                 c02 = (double)a[3] * a[7] - (double)a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double M[9];
    M[0] = c00 / det;
    M[3] = c01 / det;
    M[6] = c02 / det;
    M[1] = ((double)a[2] * a[7] - (double)a[1] * a[8]) / det;
    M[4] = ((double)a[0] * a[8] - (double)a[2] * a[6]) / det;
    M[7] = ((double)a[1] * a[6] - (double)a[0] * a[7]) / det;
    M[2] = ((double)a[1] * a[5] - (double)a[2] * a[4]) / det;
    M[5] = ((double)a[2] * a[3] - (double)a[0] * a[5]) / det;
    M[8] = ((double)a[0] * a[4] - (double)a[1] * a[3]) / det;
    const double ds[4][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}};
    double best = 0.0;
    for (int k = 0; k < 4; ++k) {
        double n2 = 0.0;
        for (int r = 0; r < 3; ++r) {
            const double v = M[r * 3] * ds[k][0] + M[r * 3 + 1] * ds[k][1] + M[r * 3 + 2] * ds[k][2];
            n2 += v * v;
        }
        best = fmax(best, sqrt(n2));
    }
    return best;
}

// Largest |value| of each view's sinogram block (one workgroup per view and
// batch element); INF when the view holds an INF or NaN.
__global__ __launch_bounds__(256) void k_view_absmax(const float* __restrict__ sino, int64_t per_view, int n_view,
                                                     float* __restrict__ view_max) {
    __shared__ float red[4];
    const int64_t base = ((int64_t)blockIdx.y * n_view + blockIdx.x) * per_view;
    float m = 0.0f;
    bool bad = false;
    for (int64_t i = threadIdx.x; i < per_view; i += blockDim.x) {
        const float a = fabsf(sino[base + i]);
        bad |= !(a <= 3.0e38f);
        m = fmaxf(m, a);
    }
    m = bad ? __builtin_huge_valf() : m;
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        view_max[(int64_t)blockIdx.y * n_view + blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// One work unit: part `part` of brick `bid` of batch element bi.  Part k of
// split takes the items of views [k * n_view / split, (k + 1) * n_view / split).
template <int MODE>
__device__ __forceinline__ void brick_unit(
    const float* __restrict__ sino, float* __restrict__ out, int64_t n_ray, int64_t vol_elems, Dims3 dims,
    BrickGrid bg, int n_view, int rows, int cols, const int32_t* __restrict__ plan,
    const uint32_t* __restrict__ totals, const float* __restrict__ view_e, const float* __restrict__ frames,
    const float* __restrict__ src,
    const float* __restrict__ dst, const float* __restrict__ minv, const float* __restrict__ b,
    const float* __restrict__ view_max, int accumulate, int split, int bid, int part, int64_t bi,
    unsigned long long* acc, uint16_t* queue, uint32_t* bcnt, float* s_red, int4* s_plan, float* s_frame) {
    const int tid = threadIdx.x;
    const Window<3> win = brick_window(dims, bg, bid);
    for (int i = tid; i < kBrickLds; i += kBrickThreads) acc[i] = 0ull;

    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const float* __restrict__ sn = sino + bi * n_ray;
    const int32_t* __restrict__ p = plan + (int64_t)bid * n_view * kPlanInts;
    const uint32_t total = totals[bid];
    const int v_lo = (int)((int64_t)part * n_view / split), v_hi = (int)((int64_t)(part + 1) * n_view / split);
    const uint32_t item_lo = (v_lo < n_view) ? (uint32_t)p[(int64_t)v_lo * kPlanInts + 3] : total;
    const uint32_t item_hi = (v_hi < n_view) ? (uint32_t)p[(int64_t)v_hi * kPlanInts + 3] : total;
    float* __restrict__ o = out + bi * vol_elems;
    const int ey = win.ext[1], ez = win.ext[2];
    const int nvox = win.ext[0] * ey * ez;

    // 1. a bound on |value| over the brick's rays: the largest per-view
    // maximum (k_view_absmax, INF for a view holding INF or NaN) over the views
    // whose shadow rectangle is not empty; and the largest midpoint error bound
    // over the views (INF when one is not a finite number)
    float vmax = 0.0f, emax = 0.0f;
    bool nonfinite = false;
    {
        const float* __restrict__ vm = view_max + bi * n_view;
        for (int v = tid; v < n_view; v += kBrickThreads) {
            if (p[(int64_t)v * kPlanInts + 2] > 0) {
                const float m = vm[v];
                nonfinite |= !(m <= 3.0e38f);
                vmax = fmaxf(vmax, m);
            }
            const float ev = view_e[v];
            emax = fmaxf(emax, (ev >= 0.0f && ev <= 3.0e38f) ? ev : __builtin_huge_valf());
        }
    }
    vmax = nonfinite ? __builtin_huge_valf() : vmax;
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
        vmax = fmaxf(vmax, __shfl_xor(vmax, o2));
        emax = fmaxf(emax, __shfl_xor(emax, o2));
    }
    if (tid % 64 == 0) {
        s_red[tid / 64] = vmax;
        reinterpret_cast<float*>(queue)[tid / 64] = emax;  // the queue is free until the items start
    }
    __syncthreads();
    vmax = 0.0f;
    emax = 0.0f;
    for (int k = 0; k < kBrickThreads / 64; ++k) {
        vmax = fmaxf(vmax, s_red[k]);
        emax = fmaxf(emax, reinterpret_cast<const float*>(queue)[k]);
    }
    // Both are block-uniform: keep them in scalar registers.
    vmax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vmax)));
    emax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(emax)));
    // 2^-10 of slack for the traversal's own rounding of the voxel coordinates
    const float e_brick = emax * 1.0009765625f;
    __syncthreads();  // queue is reused below

    // 2. power-of-two scale 2^F.  A ray's segments in one voxel add up to at
    // most the voxel diagonal, so |sum| <= total * vmax * diag per voxel; the
    // factor 2 covers the rounding of len and of the product.  2^F maps that
    // bound below 2^62 (no overflow of the signed 64-bit sums) and each single
    // contribution below 2^50 (the range of the conversion in to_fixed).  It is
    // applied to each sinogram value as two float factors, so any F in
    // [-250, 250] is exact.  The absolute error of a voxel is at most half a
    // quantum 2^-F per contribution.  Non-finite data, or products that could
    // overflow a float, take float atomics into global memory instead (INF/NaN
    // propagate as with the reference's atomics).  (fixed_scale in
    // fixed_point.h is the same computation; it is written out here because
    // the brick kernel sits at its register budget and the call's different
    // instruction order costs it spills.)
    const double cmax = 2.0 * (double)vmax * voxel_diagonal(g);
    const bool fixed = g.finite && cmax <= 1.0e38;
    const double bound = cmax * (double)(total > 0 ? total : 1);
    const int e = (bound > 0.0) ? ilogb(bound) + 1 : 0;  // bound < 2^e
    const int ec = (cmax > 0.0) ? ilogb(cmax) + 1 : 0;   // cmax < 2^ec
    const int F = fixed ? max(-250, min(min(62 - e, 50 - ec), 250)) : 0;
    const float s1 = ldexpf(1.0f, F / 2), s2 = ldexpf(1.0f, F - F / 2);  // 2^F as two normal floats

    // 3. trace and accumulate
    if (vmax != 0.0f) {
        if (fixed) {
            auto add = [&](uint32_t li, float v, float len) { atomicAdd(&acc[li], to_fixed(v * len)); };
            if (g.diag)
                brick_items<MODE, true>(g, win, sn, src, dst, p, frames, n_view, rows, cols, item_lo, item_hi, total, e_brick, s1,
                                        s2, queue, bcnt, s_plan, s_frame, add);
            else
                brick_items<MODE, false>(g, win, sn, src, dst, p, frames, n_view, rows, cols, item_lo, item_hi, total, e_brick, s1,
                                         s2, queue, bcnt, s_plan, s_frame, add);
        } else {
            if (split == 1 && !accumulate) {  // split > 1: the host zeroed out
                for (int i = tid; i < nvox; i += kBrickThreads) {
                    const int k = i % ez, j = (i / ez) % ey, ii = i / (ez * ey);
                    o[((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k)] =
                        0.0f;
                }
            }
            __syncthreads();
            auto add = [&](uint32_t li, float v, float len) {
                const uint32_t r = li % kStrideX;
                const int ii = (int)(li / kStrideX), j = (int)(r / kStrideY), k = (int)(r % kStrideY);
                const int64_t gi =
                    ((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k);
                unsafeAtomicAdd(o + gi, v * len);
            };
            if (g.diag)
                brick_items<MODE, true>(g, win, sn, src, dst, p, frames, n_view, rows, cols, item_lo, item_hi, total, e_brick, s1,
                                        s2, queue, bcnt, s_plan, s_frame, add);
            else
                brick_items<MODE, false>(g, win, sn, src, dst, p, frames, n_view, rows, cols, item_lo, item_hi, total, e_brick, s1,
                                         s2, queue, bcnt, s_plan, s_frame, add);
        }
    } else if (!fixed && !accumulate && split == 1) {
        for (int i = tid; i < nvox; i += kBrickThreads) {
            const int k = i % ez, j = (i / ez) % ey, ii = i / (ez * ey);
            o[((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k)] = 0.0f;
        }
    }
    if (!fixed || (split > 1 && vmax == 0.0f)) return;
    __syncthreads();

    // 4. write the brick (k fastest: coalesced rows of the volume).  With two
    // parts each adds its rounded sum to the zeroed output: 0 + a + b rounds
    // the same in either order, so the result does not depend on which part
    // finishes first.
    const double inv = ldexp(1.0, -F);
    for (int i = tid; i < nvox; i += kBrickThreads) {
        const int k = i % ez, j = (i / ez) % ey, ii = i / (ez * ey);
        const int64_t gi = ((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k);
        const float v = (float)((double)(long long)acc[ii * kStrideX + j * kStrideY + k] * inv);
        if (split > 1)
            add_f32_ieee(o + gi, v);
        else
            o[gi] = accumulate ? o[gi] + v : v;
    }
}

// One workgroup per work unit (x) and batch element (y).  Consecutive units go
// round-robin to the 8 XCDs, which keeps the XCDs' loads even (a contiguous
// run of bricks per XCD would give the XCDs holding the centre of the volume
// the most work).
template <int MODE>
__global__ __launch_bounds__(kBrickThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_back_brick(
    const float* __restrict__ sino, float* __restrict__ out, int64_t n_ray, int64_t vol_elems, Dims3 dims,
    BrickGrid bg, int n_view, int rows, int cols, const int32_t* __restrict__ plan,
    const uint32_t* __restrict__ totals, const float* __restrict__ view_e, const float* __restrict__ frames,
    const float* __restrict__ src,
    const float* __restrict__ dst, const float* __restrict__ minv, const float* __restrict__ b,
    const float* __restrict__ view_max, int accumulate, int split) {
    __shared__ unsigned long long acc[kBrickLds];
    __shared__ __align__(16) uint16_t queue[kChunk];
    __shared__ uint32_t bcnt[kBuckets + 1];
    __shared__ float s_red[kBrickThreads / 64];
    __shared__ int4 s_plan[kStage + 1];
    __shared__ float s_frame[kStage * kFrameFloats];
    const uint32_t unit = blockIdx.x;
    brick_unit<MODE>(sino, out, n_ray, vol_elems, dims, bg, n_view, rows, cols, plan, totals, view_e, frames, src, dst,
                     minv, b, view_max, accumulate, split, (int)(unit / (uint32_t)split),
                     (int)(unit % (uint32_t)split), (int64_t)blockIdx.y, acc, queue, bcnt, s_red, s_plan, s_frame);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int ctlab_cone_frames_3d(const float* src, const float* dst, int64_t n_view, int64_t det_rows,
                                    int64_t det_cols, const float* M_inv, const float* b, int64_t n_x, int64_t n_y,
                                    int64_t n_z, float* frames, ctlab_stream_t stream) {
    const char* fn = "cone_frames_3d";
    const int64_t n[3] = {n_x, n_y, n_z};
    for (int a = 0; a < 3; ++a) CTLAB_REQUIRE(n[a] >= 1 && n[a] < (1LL << 24), "%s: bad grid dimension", fn);
    CTLAB_REQUIRE(n_view >= 1 && n_view <= 0x7fffffffLL, "%s: n_view %lld", fn, (long long)n_view);
    CTLAB_REQUIRE(det_rows >= 1 && det_cols >= 1 && det_rows * det_cols <= 0x7fffffffLL, "%s: detector %lld x %lld",
                  fn, (long long)det_rows, (long long)det_cols);
    CTLAB_REQUIRE(src && dst && M_inv && b && frames, "%s: null pointer", fn);
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    hipLaunchKernelGGL(k_cone_frames, dim3((unsigned)n_view), dim3(256), 0, as_stream(stream), src, dst,
                       (int)det_rows, (int)det_cols, d, M_inv, b, frames);
    return check_launch(fn);
}

static BrickGrid brick_grid(const int64_t* n) {
    BrickGrid bg;
    bg.nbx = (int)((n[0] + kBrickX - 1) / kBrickX);
    bg.nby = (int)((n[1] + kBrickY - 1) / kBrickY);
    bg.nbz = (int)((n[2] + kBrickZ - 1) / kBrickZ);
    return bg;
}

extern "C" int64_t ctlab_brick_plan_ints_3d(int64_t n_x, int64_t n_y, int64_t n_z, int64_t n_view) {
    const int64_t n[3] = {n_x, n_y, n_z};
    if (n_x < 1 || n_y < 1 || n_z < 1 || n_view < 0) return -1;
    const BrickGrid bg = brick_grid(n);
    return (int64_t)bg.nbx * bg.nby * bg.nbz * (n_view * kPlanInts + 1) + n_view * (1 + kFrameFloats);
}

extern "C" int ctlab_brick_plan_3d(const float* frames, int64_t n_view, int64_t det_rows, int64_t det_cols,
                                   int64_t n_x, int64_t n_y, int64_t n_z, int32_t* plan, ctlab_stream_t stream) {
    const char* fn = "brick_plan_3d";
    const int64_t n[3] = {n_x, n_y, n_z};
    for (int a = 0; a < 3; ++a) CTLAB_REQUIRE(n[a] >= 1 && n[a] < (1LL << 24), "%s: bad grid dimension", fn);
    CTLAB_REQUIRE(n_view >= 1 && det_rows >= 1 && det_cols >= 1, "%s: bad detector layout", fn);
    CTLAB_REQUIRE(n_view * det_rows * det_cols <= 0xffffffffLL - 4096,
                  "%s: n_view*rows*cols must be below 2^32 - 4096 for the brick plan", fn);
    CTLAB_REQUIRE(frames && plan, "%s: null pointer", fn);
    const BrickGrid bg = brick_grid(n);
    const int64_t nb = (int64_t)bg.nbx * bg.nby * bg.nbz;
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    int32_t* table = plan;
    uint32_t* totals = reinterpret_cast<uint32_t*>(plan + nb * n_view * kPlanInts);
    float* view_e = reinterpret_cast<float*>(plan + nb * (n_view * kPlanInts + 1));
    float* frames_copy = view_e + n_view;
    hipLaunchKernelGGL(k_brick_plan, dim3((unsigned)nb), dim3(256), 0, as_stream(stream), d, bg, (int)n_view,
                       (int)det_rows, (int)det_cols, frames, table, totals, view_e, frames_copy);
    return check_launch(fn);
}

namespace ctlab {

// Shared by both brick entry points: checks, per-view maxima into view_max
// (batch * n_view floats, every entry written), the zero fill and the launch.
static int back_brick(const char* fn, const float* sino, int64_t batch, const float* src, const float* dst,
                      int64_t n_view, int64_t det_rows, int64_t det_cols, const int32_t* plan, const float* M_inv,
                      const float* b, int64_t n_x, int64_t n_y, int64_t n_z, int mode, int accumulate, float* out,
                      float* view_max, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    int64_t V = 1;
    for (int a = 0; a < 3; ++a) {
        CTLAB_REQUIRE(n[a] >= 1, "%s: bad grid dimension", fn);
        V *= n[a];
    }
    CTLAB_REQUIRE(V <= kMaxVoxels, "%s: grid has more than 2^31-1 voxels", fn);
    int st = check_index_limits(fn, n, 3);
    if (st) return st;
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range", fn, (long long)batch);
    CTLAB_REQUIRE(mode == CTLAB_MODE_TORCH || mode == CTLAB_MODE_OTF, "%s: unknown mode %d", fn, mode);
    CTLAB_REQUIRE(n_view >= 1 && det_rows >= 1 && det_cols >= 1 &&
                      n_view * det_rows * det_cols <= 0xffffffffLL - 4096,
                  "%s: bad detector layout (n_view*rows*cols must be below 2^32 - 4096)", fn);
    CTLAB_REQUIRE(out && sino && src && dst && plan && M_inv && b && view_max, "%s: null pointer", fn);
    const int64_t n_ray = n_view * det_rows * det_cols;
    const BrickGrid bg = brick_grid(n);
    const int64_t nb = (int64_t)bg.nbx * bg.nby * bg.nbz;
    const uint32_t* totals = reinterpret_cast<const uint32_t*>(plan + nb * n_view * kPlanInts);
    const float* view_e = reinterpret_cast<const float*>(plan + nb * (n_view * kPlanInts + 1));
    const float* frames = view_e + n_view;
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    // Two work units per brick (halves of the views) halve the longest unit and
    // with it the tail of the launch; they add into a zeroed output, which the
    // accumulate mode cannot use.
    const int split = accumulate ? 1 : 2;
    dim3 grid((unsigned)(nb * split), (unsigned)batch);
    hipLaunchKernelGGL(k_view_absmax, dim3((unsigned)n_view, (unsigned)batch), dim3(256), 0, as_stream(stream), sino,
                       det_rows * det_cols, (int)n_view, view_max);
    st = check_launch(fn);
    if (st) return st;
    if (split > 1) {
        st = fill_f32(out, batch * V, 0.0f, as_stream(stream));
        if (st) return st;
    }
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL(k_back_brick<MODE_TORCH>, grid, dim3(kBrickThreads), 0, as_stream(stream), sino, out,
                           n_ray, V, d, bg, (int)n_view, (int)det_rows, (int)det_cols, plan, totals, view_e, frames, src,
                           dst, M_inv, b, view_max, accumulate, split);
    else
        hipLaunchKernelGGL(k_back_brick<MODE_OTF>, grid, dim3(kBrickThreads), 0, as_stream(stream), sino, out,
                           n_ray, V, d, bg, (int)n_view, (int)det_rows, (int)det_cols, plan, totals, view_e, frames, src,
                           dst, M_inv, b, view_max, accumulate, split);
    return check_launch(fn);
}

// Per-view maxima of the entry point that allocates nothing.  In host code the
// name `scratch` is only a stand-in: kernels and copies get the device address
// that hipGetSymbolAddress returns for the current device, kept per device.
static __device__ __align__(256) float scratch[CTLAB_BRICK_SCRATCH_FLOATS];

constexpr int kMaxDevices = 64;
static float* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_lock;

// Device address of this device's copy of `scratch` (looked up once per device,
// with that device current).
static int scratch_of_current_device(const char* fn, float** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(CTLAB_EHIP, "%s: hipGetDevice: %s", fn, hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(CTLAB_EINVAL, "%s: device %d beyond %d", fn, dev, kMaxDevices);
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CTLAB_EHIP, "%s: hipGetSymbolAddress(scratch): %s", fn, hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return CTLAB_OK;
}

}  // namespace ctlab

extern "C" int ctlab_back_project_3d_brick(const float* sino, int64_t batch, const float* src, const float* dst,
                                           int64_t n_view, int64_t det_rows, int64_t det_cols, const int32_t* plan,
                                           const float* M_inv, const float* b, int64_t n_x, int64_t n_y,
                                           int64_t n_z, int mode, int accumulate, float* out, float* view_max,
                                           ctlab_stream_t stream) {
    return ctlab::back_brick("back_project_3d_brick", sino, batch, src, dst, n_view, det_rows, det_cols, plan, M_inv,
                             b, n_x, n_y, n_z, mode, accumulate, out, view_max, stream);
}

extern "C" int ctlab_back_project_3d_brick_noalloc(const float* sino, int64_t batch, const float* src,
                                                   const float* dst, int64_t n_view, int64_t det_rows,
                                                   int64_t det_cols, const int32_t* plan, const float* M_inv,
                                                   const float* b, int64_t n_x, int64_t n_y, int64_t n_z, int mode,
                                                   int accumulate, float* out, ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "back_project_3d_brick_noalloc";
    CTLAB_REQUIRE(batch >= 1 && n_view >= 1 && batch <= CTLAB_BRICK_SCRATCH_FLOATS / n_view,
                  "%s: batch * n_view = %lld * %lld exceeds the %d per-view maxima of the static workspace", fn,
                  (long long)batch, (long long)n_view, (int)CTLAB_BRICK_SCRATCH_FLOATS);
    float* view_max = nullptr;
    const int st = scratch_of_current_device(fn, &view_max);
    if (st) return st;
    return back_brick(fn, sino, batch, src, dst, n_view, det_rows, det_cols, plan, M_inv, b, n_x, n_y, n_z, mode,
                      accumulate, out, view_max, stream);
}

// ===========================================================================
// Rays in any order: brick lists
// ===========================================================================
//
// Without a detector layout a brick's rays cannot be read off a shadow
// rectangle.  A binning pass lists instead, for every brick, the rays that can
// have a segment in it: those passing through the brick's cells widened by
// E + 0.05 voxel, with E the ray's own midpoint error bound, which is the test
// the filing pass above applies with a bound over views.  A list variant of the
// brick kernel then traces the rays of each list exactly as the shadow variant
// traces the rays of its rectangles.  The lists depend only on the geometry
// (rays, grid, map), so they are built once per ray set.

namespace ctlab {

constexpr int kBinThreads = 256;
constexpr int64_t kBinRaysPerBlock = 8192;   // rays per block of the fill pass
constexpr int kBinMaxLdsBricks = 16384;      // bricks counted in LDS (64 KB per block)

__device__ __forceinline__ int floor_div_clamped(float c, int B, int nb) {
    const float q = floorf(c / (float)B);
    return (int)fminf(fmaxf(q, 0.0f), (float)(nb - 1));
}

// Calls emit(brick) for every brick whose cells, widened by m = E + 0.05, the
// ray's segment t in [0, 1] passes through (slab_estimate > 0), E being the
// ray's own midpoint error bound as trace_ray_window computes it.  Returns E,
// or -1 for a ray the traversal gives no segment (non-finite input, E above 8,
// or a zero-length ray under the otf rule).  Candidates are enumerated layer by
// layer along the axis the ray crosses most bricks of; within a layer (widened
// by m + 1 voxel) the other two axes span one or two bricks each.
template <class Emit>
__device__ float ray_bricks(const Grid<3>& g, const Dims3& dims, const BrickGrid& bg, int mode,
                            const float* __restrict__ src, const float* __restrict__ dst, Emit&& emit) {
    if (!g.finite) return -1.0f;
    bool fin = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) fin = fin && (fabsf(src[a]) <= 3.0e38f) && (fabsf(dst[a]) <= 3.0e38f);
    if (!fin) return -1.0f;
    if (mode == MODE_OTF) {
        const float v0 = dst[0] - src[0], v1 = dst[1] - src[1], v2 = dst[2] - src[2];
        const float rl = __builtin_sqrtf(__builtin_fmaf(v2, v2, __builtin_fmaf(v1, v1, v0 * v0)));
        if (rl < 1.0e-12f) return -1.0f;
    }
    float sv[3], dv[3];
    to_voxel<3, MODE_TORCH>(g, src, sv);
    to_voxel<3, MODE_TORCH>(g, dst, dv);
    float E = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float mag = fabsf(sv[a]) + fabsf(dv[a]) + (float)g.n[a] + 1.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j) mag += fabsf(g.minv[a * 3 + j]) * (fabsf(src[j]) + fabsf(dst[j]) + fabsf(g.b[j]));
        const float Ea = mag * 7.62939453125e-06f;  // 2^-17
        if (!(Ea <= 8.0f)) return -1.0f;
        E = fmaxf(E, Ea);
    }
    const float m = E + 0.05f, w = m + 1.0f;
    const int B[3] = {kBrickX, kBrickY, kBrickZ};
    const int nb[3] = {bg.nbx, bg.nby, bg.nbz};
    float dir[3];
    float t0 = 0.0f, t1 = 1.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        dir[a] = dv[a] - sv[a];
        const float lo = -0.5f - w, hi = (float)dims.n[a] - 0.5f + w;
        if (dir[a] == 0.0f) {
            if (!(sv[a] >= lo && sv[a] <= hi)) return E;
            continue;
        }
        const float ta = (lo - sv[a]) / dir[a], tb = (hi - sv[a]) / dir[a];
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    if (!(t0 <= t1)) return E;
    int D = 0;
    float best = -1.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float r = fabsf(dir[a]) / (float)B[a];
        if (r > best) {
            best = r;
            D = a;
        }
    }
    const int A1 = (D == 0) ? 1 : 0, A2 = (D == 2) ? 1 : 2;
    const float cA = sv[D] + t0 * dir[D], cB = sv[D] + t1 * dir[D];
    const int k0 = floor_div_clamped(fminf(cA, cB) + 0.5f - w, B[D], nb[D]);
    const int k1 = floor_div_clamped(fmaxf(cA, cB) + 0.5f + w, B[D], nb[D]);
    for (int k = k0; k <= k1; ++k) {
        float ta = t0, tb = t1;
        if (dir[D] != 0.0f) {
            const float lo = (float)(k * B[D]) - 0.5f - w, hi = (float)((k + 1) * B[D]) - 0.5f + w;
            const float x0 = (lo - sv[D]) / dir[D], x1 = (hi - sv[D]) / dir[D];
            ta = fmaxf(t0, fminf(x0, x1));
            tb = fminf(t1, fmaxf(x0, x1));
            if (!(ta <= tb)) continue;
        }
        int j0[2], j1[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int a = q ? A2 : A1;
            const float ca = sv[a] + ta * dir[a], cb = sv[a] + tb * dir[a];
            j0[q] = floor_div_clamped(fminf(ca, cb) + 0.5f - w, B[a], nb[a]);
            j1[q] = floor_div_clamped(fmaxf(ca, cb) + 0.5f + w, B[a], nb[a]);
        }
        for (int i1 = j0[0]; i1 <= j1[0]; ++i1) {
            for (int i2 = j0[1]; i2 <= j1[1]; ++i2) {
                int bi[3];
                bi[D] = k;
                bi[A1] = i1;
                bi[A2] = i2;
                float blo[3], bhi[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int lo = bi[a] * B[a];
                    const int ext = min(B[a], dims.n[a] - lo);
                    blo[a] = (float)lo - 0.5f - m;
                    bhi[a] = (float)(lo + ext) - 0.5f + m;
                }
                if (slab_estimate(sv, dv, blo, bhi, 0.0f))
                    emit((uint32_t)((bi[0] * nb[1] + bi[1]) * nb[2] + bi[2]));
            }
        }
    }
    return E;
}

template <int ND>
__device__ __forceinline__ void load_ray3(const float* __restrict__ src, const float* __restrict__ dst, int64_t r,
                                          float* s, float* d) {
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        s[a] = src[3 * r + a];
        d[a] = dst[3 * r + a];
    }
}

// Pass 1: the number of rays per brick (LDS histogram per block when the grid
// has few enough bricks) and the largest E over the rays with segments.
template <bool LDS>
__global__ __launch_bounds__(kBinThreads) void k_bin_count(const float* __restrict__ src,
                                                            const float* __restrict__ dst, int64_t n_ray,
                                                            Dims3 dims, BrickGrid bg,
                                                            const float* __restrict__ minv,
                                                            const float* __restrict__ b, int mode,
                                                            uint32_t* __restrict__ counts,
                                                            unsigned int* __restrict__ emax_bits) {
    extern __shared__ uint32_t hist[];
    const int nbk = bg.nbx * bg.nby * bg.nbz;
    if (LDS) {
        for (int i = threadIdx.x; i < nbk; i += kBinThreads) hist[i] = 0u;
        __syncthreads();
    }
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    float emax = 0.0f;
    for (int64_t r = (int64_t)blockIdx.x * kBinThreads + threadIdx.x; r < n_ray;
         r += (int64_t)gridDim.x * kBinThreads) {
        float s[3], d[3];
        load_ray3<3>(src, dst, r, s, d);
        const float e = ray_bricks(g, dims, bg, mode, s, d, [&](uint32_t bk) {
            if (LDS)
                atomicAdd(&hist[bk], 1u);
            else
                atomicAdd(&counts[bk], 1u);
        });
        emax = fmaxf(emax, e);
    }
    for (int o = 32; o > 0; o >>= 1) emax = fmaxf(emax, __shfl_xor(emax, o));
    if (threadIdx.x % 64 == 0) atomicMax(emax_bits, __float_as_uint(emax));  // emax >= 0: bits order as values
    if (LDS) {
        __syncthreads();
        for (int i = threadIdx.x; i < nbk; i += kBinThreads)
            if (hist[i]) atomicAdd(&counts[i], hist[i]);
    }
}

// Pass 2: each block lists the rays of its range.  It counts its rays per
// brick in LDS, reserves one contiguous run per brick with a single global
// atomic, and places the ray ids with LDS cursors.  The order of ids within a
// brick's list depends on timing; the sums the brick kernel forms from them do
// not (fixed point).
template <bool LDS>
__global__ __launch_bounds__(kBinThreads) void k_bin_fill(const float* __restrict__ src,
                                                           const float* __restrict__ dst, int64_t n_ray,
                                                           Dims3 dims, BrickGrid bg,
                                                           const float* __restrict__ minv,
                                                           const float* __restrict__ b, int mode,
                                                           uint32_t* __restrict__ cursor,
                                                           uint32_t* __restrict__ ids, int64_t n_ids) {
    extern __shared__ uint32_t loc[];
    const int nbk = bg.nbx * bg.nby * bg.nbz;
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const int64_t r0 = (int64_t)blockIdx.x * kBinRaysPerBlock;
    const int64_t r1 = min(n_ray, r0 + kBinRaysPerBlock);
    if (LDS) {
        for (int i = threadIdx.x; i < nbk; i += kBinThreads) loc[i] = 0u;
        __syncthreads();
        for (int64_t r = r0 + threadIdx.x; r < r1; r += kBinThreads) {
            float s[3], d[3];
            load_ray3<3>(src, dst, r, s, d);
            ray_bricks(g, dims, bg, mode, s, d, [&](uint32_t bk) { atomicAdd(&loc[bk], 1u); });
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nbk; i += kBinThreads) {
            const uint32_t c = loc[i];
            if (c) loc[i] = atomicAdd(&cursor[i], c);
        }
        __syncthreads();
        for (int64_t r = r0 + threadIdx.x; r < r1; r += kBinThreads) {
            float s[3], d[3];
            load_ray3<3>(src, dst, r, s, d);
            ray_bricks(g, dims, bg, mode, s, d, [&](uint32_t bk) {
                const uint32_t slot = atomicAdd(&loc[bk], 1u);
                if ((int64_t)slot < n_ids) ids[slot] = (uint32_t)r;
            });
        }
    } else {
        for (int64_t r = r0 + threadIdx.x; r < r1; r += kBinThreads) {
            float s[3], d[3];
            load_ray3<3>(src, dst, r, s, d);
            ray_bricks(g, dims, bg, mode, s, d, [&](uint32_t bk) {
                const uint32_t slot = atomicAdd(&cursor[bk], 1u);
                if ((int64_t)slot < n_ids) ids[slot] = (uint32_t)r;
            });
        }
    }
}

// Largest |value| of each batch element's sinogram (INF when one is INF or
// NaN), as the bits of a non-negative float: max_bits must hold zeros.
__global__ __launch_bounds__(256) void k_absmax_all(const float* __restrict__ sino, int64_t n_ray,
                                                    unsigned int* __restrict__ max_bits) {
    const float* __restrict__ s = sino + (int64_t)blockIdx.y * n_ray;
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n_ray; i += (int64_t)gridDim.x * 256) {
        const float a = fabsf(s[i]);
        m = fmaxf(m, (a <= 3.0e38f) ? a : __builtin_huge_valf());
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (threadIdx.x % 64 == 0) atomicMax(max_bits + blockIdx.y, __float_as_uint(m));
}

static int check_grid3(const char* fn, const int64_t* n) {
    int64_t V = 1;
    for (int a = 0; a < 3; ++a) {
        CTLAB_REQUIRE(n[a] >= 1, "%s: bad grid dimension", fn);
        V *= n[a];
    }
    CTLAB_REQUIRE(V <= kMaxVoxels, "%s: grid has more than 2^31-1 voxels", fn);
    return check_index_limits(fn, n, 3);
}

}  // namespace ctlab

extern "C" int64_t ctlab_brick_count_3d(int64_t n_x, int64_t n_y, int64_t n_z) {
    using namespace ctlab;
    const int64_t n[3] = {n_x, n_y, n_z};
    if (n_x < 1 || n_y < 1 || n_z < 1) return -1;
    const BrickGrid bg = brick_grid(n);
    return (int64_t)bg.nbx * bg.nby * bg.nbz;
}

extern "C" int ctlab_brick_bins_count_3d(const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                         const float* b, int64_t n_x, int64_t n_y, int64_t n_z, int mode,
                                         uint32_t* counts, float* emax, ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "brick_bins_count_3d";
    const int64_t n[3] = {n_x, n_y, n_z};
    int st = check_grid3(fn, n);
    if (st) return st;
    CTLAB_REQUIRE(mode == CTLAB_MODE_TORCH || mode == CTLAB_MODE_OTF, "%s: unknown mode %d", fn, mode);
    CTLAB_REQUIRE(n_ray >= 0 && n_ray <= 0xffffffffLL, "%s: n_ray %lld out of range", fn, (long long)n_ray);
    CTLAB_REQUIRE(counts && emax && (n_ray == 0 || (src && dst && M_inv && b)), "%s: null pointer", fn);
    const BrickGrid bg = brick_grid(n);
    const int64_t nbk = (int64_t)bg.nbx * bg.nby * bg.nbz;
    st = fill_f32(reinterpret_cast<float*>(counts), nbk, 0.0f, as_stream(stream));  // zero bits
    if (st) return st;
    st = fill_f32(emax, 1, 0.0f, as_stream(stream));
    if (st || n_ray == 0) return st;
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    const unsigned blocks = (unsigned)std::min<int64_t>((n_ray + kBinThreads - 1) / kBinThreads, 4096);
    if (nbk <= kBinMaxLdsBricks)
        hipLaunchKernelGGL(k_bin_count<true>, dim3(blocks), dim3(kBinThreads), (size_t)nbk * 4, as_stream(stream), src,
                           dst, n_ray, d, bg, M_inv, b, mode, counts, reinterpret_cast<unsigned int*>(emax));
    else
        hipLaunchKernelGGL(k_bin_count<false>, dim3(blocks), dim3(kBinThreads), 0, as_stream(stream), src, dst, n_ray,
                           d, bg, M_inv, b, mode, counts, reinterpret_cast<unsigned int*>(emax));
    return check_launch(fn);
}

extern "C" int ctlab_brick_bins_fill_3d(const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                        const float* b, int64_t n_x, int64_t n_y, int64_t n_z, int mode,
                                        uint32_t* cursor, uint32_t* ids, int64_t n_ids, ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "brick_bins_fill_3d";
    const int64_t n[3] = {n_x, n_y, n_z};
    int st = check_grid3(fn, n);
    if (st) return st;
    CTLAB_REQUIRE(mode == CTLAB_MODE_TORCH || mode == CTLAB_MODE_OTF, "%s: unknown mode %d", fn, mode);
    CTLAB_REQUIRE(n_ray >= 0 && n_ray <= 0xffffffffLL, "%s: n_ray %lld out of range", fn, (long long)n_ray);
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(cursor && ids && src && dst && M_inv && b && n_ids >= 0, "%s: null pointer", fn);
    const BrickGrid bg = brick_grid(n);
    const int64_t nbk = (int64_t)bg.nbx * bg.nby * bg.nbz;
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    const unsigned blocks = (unsigned)((n_ray + kBinRaysPerBlock - 1) / kBinRaysPerBlock);
    if (nbk <= kBinMaxLdsBricks)
        hipLaunchKernelGGL(k_bin_fill<true>, dim3(blocks), dim3(kBinThreads), (size_t)nbk * 4, as_stream(stream), src,
                           dst, n_ray, d, bg, M_inv, b, mode, cursor, ids, n_ids);
    else
        hipLaunchKernelGGL(k_bin_fill<false>, dim3(blocks), dim3(kBinThreads), 0, as_stream(stream), src, dst, n_ray, d,
                           bg, M_inv, b, mode, cursor, ids, n_ids);
    return check_launch(fn);
}

// ===========================================================================
// Filed unit lists: the shadow variant's filing done once per geometry
// ===========================================================================
//
// Which rays of a brick's shadow rectangles pass through the brick's widened
// cells, and how many crossings each has there, depends only on the geometry.
// The shadow variant files them again in every call, chunk by chunk, with a
// block barrier per chunk.  Here one pass per geometry lists, for every work
// unit (a brick and one half of the views, the shadow variant's units), the
// rays that pass the same slab test, grouped by the same quarter-octave
// estimate of their crossing count, longest class first.  The back projector
// over these lists (k_back_filed) has each wave take the next 64 rays of its
// unit from an LDS counter and trace them: no filing, no item decoding and no
// barrier until the unit is done.  Each unit adds up the segments of the same
// rays as the shadow variant's unit, in 64-bit fixed point with the same
// scale, so the output is the same bit for bit.

namespace ctlab {

constexpr int kFileThreads = 256;
constexpr int kFileHalves = 2;     // units per brick: the first and the second half of the views
constexpr int kFileBuckets = 64;    // eighth-octave length classes of the filed unit lists
constexpr int kFileMaxChunks = 64;  // view chunks per unit, at most

// A unit's list is cut into chunks of consecutive views (at least 8 views, at
// most kFileMaxChunks chunks per unit), each chunk's rays grouped by length
// class.  Waves take the list in order, so the workgroups of the launch move
// through the views together and read the same rays while they sit in L2.
// Measured on the 256^3 / 720-view cone beam (dense back projection): one
// chunk per unit 28.9 ms, 32 views per chunk 27.3, 16 views 26.8, 8 views 26.6.
static int filed_chunk_views(int64_t n_view) {
    const int64_t half = (n_view + 1) / 2;
    return (int)std::max<int64_t>(8, (half + kFileMaxChunks - 1) / kFileMaxChunks);
}

static int filed_unit_slots(int64_t n_view) {
    const int64_t half = (n_view + 1) / 2, V = filed_chunk_views(n_view);
    return (int)((half + V - 1) / V) * kFileBuckets;
}
constexpr int kFileWaveQueue = 128;  // per-wave LDS queue of (ray, value) with a nonzero value
// Back projector workgroups: 12 waves, three per SIMD, two workgroups per CU
// (6 waves per SIMD, 80 VGPRs each).  Wave counts that are not a multiple of
// four leave the SIMDs unevenly loaded; measured on the 256^3 / 720-view
// cone beam: 768 threads 29.4-30.9 ms, 1024 (64 VGPRs, more spills) 33.9-34.8,
// 512 31.5, 640 41.4, 896 39.4.
constexpr int kFiledThreads = 768;

// Largest midpoint error bound over all views (INF when one is not a finite
// number), times the 2^-10 slack of brick_unit: the bound the shadow variant
// widens its bricks by.
__device__ __forceinline__ float filed_brick_e(const float* __restrict__ view_e, int n_view, float* red) {
    float emax = 0.0f;
    for (int v = threadIdx.x; v < n_view; v += blockDim.x) {
        const float ev = view_e[v];
        emax = fmaxf(emax, (ev >= 0.0f && ev <= 3.0e38f) ? ev : __builtin_huge_valf());
    }
    return block_max(emax, red) * 1.0009765625f;
}

// Pass 1 (FILL false) counts the rays of every unit per view chunk and length
// class into counts[unit * S + chunk * kFileBuckets + j], j = kFileBuckets - 1 -
// class in even chunks and class in odd ones (S = chunks * kFileBuckets); pass 2 (FILL true) writes
// each ray id at offsets[that index] + a cursor, offsets being the exclusive
// prefix sums of the counts.  One block per unit; items are decoded as
// brick_items decodes them (plan entries and frames of kStage views staged in
// LDS per chunk).
template <bool FILL>
__global__ __launch_bounds__(kFileThreads) void k_file_units(
    const float* __restrict__ src, const float* __restrict__ dst, Dims3 dims, BrickGrid bg, int n_view, int rows,
    int cols, const int32_t* __restrict__ plan, const uint32_t* __restrict__ totals,
    const float* __restrict__ view_e, const float* __restrict__ frames, const float* __restrict__ minv,
    const float* __restrict__ b, int chunk_views, int unit_slots, uint32_t* __restrict__ counts,
    const int64_t* __restrict__ offsets, uint32_t* __restrict__ ids, int64_t n_ids) {
    __shared__ uint32_t bcnt[kFileMaxChunks * kFileBuckets];
    __shared__ int4 s_plan[kStage + 1];
    __shared__ float s_frame[kStage * kFrameFloats];
    __shared__ float red[kFileThreads / 64];
    const int tid = threadIdx.x;
    const int unit = blockIdx.x, bid = unit / kFileHalves, part = unit % kFileHalves;
    const Window<3> win = brick_window(dims, bg, bid);
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const int32_t* __restrict__ p = plan + (int64_t)bid * n_view * kPlanInts;
    const uint32_t total = totals[bid];
    const int v_lo = (int)((int64_t)part * n_view / kFileHalves);
    const int v_hi = (int)((int64_t)(part + 1) * n_view / kFileHalves);
    const uint32_t item_lo = (v_lo < n_view) ? (uint32_t)p[(int64_t)v_lo * kPlanInts + 3] : total;
    const uint32_t item_hi = (v_hi < n_view) ? (uint32_t)p[(int64_t)v_hi * kPlanInts + 3] : total;
    const float e_brick = filed_brick_e(view_e, n_view, red);
    const float m = ((e_brick <= 8.0f) ? e_brick : 8.0f) + 0.05f;
    float blo[3], bhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        blo[a] = (float)win.lo[a] - 0.5f - m;
        bhi[a] = (float)(win.lo[a] + win.ext[a]) - 0.5f + m;
    }
    for (int i = tid; i < unit_slots; i += kFileThreads)
        bcnt[i] = FILL ? (uint32_t)offsets[(int64_t)unit * unit_slots + i] : 0u;
    int vs = 0;
    {
        int lo = 0, hi = n_view;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((uint32_t)p[mid * kPlanInts + 3] <= item_lo) lo = mid; else hi = mid;
        }
        vs = lo;
    }
    constexpr int kFileChunk = 4 * kFileThreads;
    uint32_t c0 = item_lo;
    while (c0 < item_hi) {
        if (tid <= kStage) {
            const int v = vs + tid;
            int4 e = make_int4(0, 0, 1, (int)total);  // past the last view: prefix = total
            if (v < n_view) e = *reinterpret_cast<const int4*>(p + (int64_t)v * kPlanInts);
            s_plan[tid] = e;
            if (tid < kStage && v < n_view) {
                const float* __restrict__ f = frames + (int64_t)v * kFrameFloats;
#pragma unroll
                for (int k = 0; k < kFrameFloats; ++k) s_frame[tid * kFrameFloats + k] = f[k];
            }
        }
        __syncthreads();
        const uint32_t lim = (uint32_t)s_plan[kStage].w;  // first item beyond the staged views
        const uint32_t c1 = min(c0 + (uint32_t)kFileChunk, min(lim, item_hi));
        for (uint32_t item = c0 + (uint32_t)tid; item < c1; item += kFileThreads) {
            const int j = staged_view(s_plan, kStage, item);
            const int4 e = s_plan[j];
            uint32_t rem;
            const int r = e.x + (int)div_rect(item - (uint32_t)e.w, (uint32_t)e.z, &rem);
            const int c = e.y + (int)rem;
            const uint32_t ray = (uint32_t)(((int64_t)(vs + j) * rows + r) * cols + c);
            const float* f = s_frame + j * kFrameFloats;
            const uint32_t est = (f[F_VALID] != 0.0f)
                                     ? frame_estimate(f, r, c, blo, bhi)
                                     : segment_estimate<MODE_TORCH>(g, src + 3ull * ray, dst + 3ull * ray, blo, bhi);
            if (est) {
                const uint32_t l2 = 31u - (uint32_t)__clz(est);
                const uint32_t bk = min(8u * l2 + ((est << 3) >> l2) - 8u, (uint32_t)(kFileBuckets - 1));
                const int ch = (vs + j - v_lo) / chunk_views;
                // classes longest first in even chunks, shortest first in odd ones: where two
                // chunks meet in the list, a wave's 64 rays have similar lengths
                const int cls = (ch & 1) ? (int)bk : kFileBuckets - 1 - (int)bk;
                const uint32_t slot = atomicAdd(&bcnt[ch * kFileBuckets + cls], 1u);
                if (FILL && (int64_t)slot < n_ids) ids[slot] = ray;
            }
        }
        vs += staged_view(s_plan, kStage + 1, c1);
        c0 = c1;
        __syncthreads();
    }
    if (!FILL)
        for (int i = tid; i < unit_slots; i += kFileThreads) counts[(int64_t)unit * unit_slots + i] = bcnt[i];
}

// The rays of list[0, n) of one unit, traced by the wave that takes them:
// each wave takes the next 64 ids from *next (LDS), queues the rays with a
// nonzero value in its own LDS queue (a zero value adds exactly zero) and
// traces 64 queued rays at a time, so its lanes stay busy on sparse sinograms.
// Every wave leaves the loop once the list is exhausted and its queue empty.
template <int MODE, bool DIAG, bool POW2, bool ZERO, class Add>
__device__ __forceinline__ void filed_items(const Grid<3>& g, const Window<3>& win, const float* __restrict__ sn,
                                            const float* __restrict__ src, const float* __restrict__ dst,
                                            const uint32_t* __restrict__ list, uint32_t n, int64_t n_ray,
                                            float e_brick, float s1, float s2, uint32_t* next, uint2* wq,
                                            Add&& add) {
    const uint32_t lane = threadIdx.x & 63u;
    const float e_hint = (e_brick <= 8.0f) ? e_brick : -1.0f;
    uint32_t pend = 0;  // queued rays (wave-uniform)
    bool more = true;   // the list has rays left (wave-uniform)
    while (true) {
        if (more && pend < 64u) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(next, 64u);
            base = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl((int)base, 0));
            if (base >= n) {
                more = false;
            } else {
                const uint32_t q = base + lane;
                uint32_t ray = 0u;
                float v = 0.0f;
                if (q < n) {
                    ray = list[q];
                    if ((int64_t)ray < n_ray) v = sn[ray];  // ids of n_ray or more: no ray
                }
                const bool nz = v != 0.0f;
                const uint64_t mask = __ballot(nz);
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                if (nz) wq[pend + below] = make_uint2(ray, __float_as_uint(v));  // pend + below < 64 + 64
                pend += (uint32_t)__popcll(mask);
                continue;
            }
        }
        if (pend == 0u) break;  // list exhausted and queue empty
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t take = min(pend, 64u);
        const uint2 item = wq[lane];
        const bool keep = 64u + lane < pend;
        const uint2 rest = keep ? wq[64u + lane] : make_uint2(0u, 0u);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (keep) wq[lane] = rest;
        pend -= take;
        if (lane < take) {
            const uint32_t ray = item.x;
            const float val = __uint_as_float(item.y) * s1 * s2;
            float s[3], d[3];
            for (int a = 0; a < 3; ++a) {
                s[a] = src[3ull * ray + a];
                d[a] = dst[3ull * ray + a];
            }
            trace_ray_window<3, MODE, DIAG, true, POW2, ZERO>(
                g, win, s, d, [&](uint32_t li, float len) { add(li, val, len); }, e_hint);
        }
    }
}

// One workgroup per unit (x) and batch element (y).  split 2: unit = (brick,
// half); split 1 (accumulate): unit = brick, both halves' lists (contiguous).
// Value bound, scale, fallback and write-back as brick_unit.
// Steps 2-4 of a filed work unit (after its value bound vmax, its midpoint
// error bound e_brick and its ray count bound total are known): the power-of-two
// scale, the traversal of list[0, n) (filed_items) and the write-back, as
// brick_unit.  acc is the zeroed LDS brick; *next is zero.
template <int MODE>
__device__ __forceinline__ void filed_unit_rest(const Grid<3>& g, const Window<3>& win, const Dims3& dims,
                                                float* __restrict__ o, const float* __restrict__ sn,
                                                const float* __restrict__ src, const float* __restrict__ dst,
                                                const uint32_t* __restrict__ list, uint32_t n, int64_t n_ray,
                                                float vmax, float e_brick, uint32_t total, int accumulate,
                                                int split, unsigned long long* acc, uint32_t* s_next,
                                                uint2* my_q) {
    const int tid = threadIdx.x;
    const int ey = win.ext[1], ez = win.ext[2];
    const int nvox = win.ext[0] * ey * ez;
    // 2. scale (as brick_unit: total bounds the number of rays of the brick)
    const double cmax = 2.0 * (double)vmax * voxel_diagonal(g);
    const bool fixed = g.finite && cmax <= 1.0e38;
    const double bound = cmax * (double)(total > 0 ? total : 1);
    const int e = (bound > 0.0) ? ilogb(bound) + 1 : 0;
    const int ec = (cmax > 0.0) ? ilogb(cmax) + 1 : 0;
    const int F = fixed ? max(-250, min(min(62 - e, 50 - ec), 250)) : 0;
    const float s1 = ldexpf(1.0f, F / 2), s2 = ldexpf(1.0f, F - F / 2);

    // 3. trace and accumulate
    if (vmax != 0.0f) {
        if (fixed) {
            // finite values: a zero-length segment adds an exact zero, so it is not tested for
            auto add = [&](uint32_t li, float v, float len) { atomicAdd(&acc[li], to_fixed(v * len)); };
            if (MODE == MODE_TORCH && g.pow2)
                filed_items<MODE, true, MODE == MODE_TORCH, true>(g, win, sn, src, dst, list, n, n_ray, e_brick,
                                                                   s1, s2, s_next, my_q, add);
            else if (g.diag)
                filed_items<MODE, true, false, true>(g, win, sn, src, dst, list, n, n_ray, e_brick, s1, s2, s_next,
                                                     my_q, add);
            else
                filed_items<MODE, false, false, true>(g, win, sn, src, dst, list, n, n_ray, e_brick, s1, s2, s_next,
                                                      my_q, add);
        } else {
            if (split == 1 && !accumulate) {
                for (int i = tid; i < nvox; i += kFiledThreads) {
                    const int k = i % ez, j = (i / ez) % ey, ii = i / (ez * ey);
                    o[((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k)] =
                        0.0f;
                }
            }
            __syncthreads();
            auto add = [&](uint32_t li, float v, float len) {
                const uint32_t r = li % kStrideX;
                const int ii = (int)(li / kStrideX), j = (int)(r / kStrideY), k = (int)(r % kStrideY);
                const int64_t gi =
                    ((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k);
                unsafeAtomicAdd(o + gi, v * len);
            };
            if (g.diag)
                filed_items<MODE, true, false, false>(g, win, sn, src, dst, list, n, n_ray, e_brick, s1, s2, s_next,
                                                      my_q, add);
            else
                filed_items<MODE, false, false, false>(g, win, sn, src, dst, list, n, n_ray, e_brick, s1, s2, s_next,
                                                       my_q, add);
        }
    } else if (!fixed && !accumulate && split == 1) {
        for (int i = tid; i < nvox; i += kFiledThreads) {
            const int k = i % ez, j = (i / ez) % ey, ii = i / (ez * ey);
            o[((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k)] = 0.0f;
        }
    }
    if (!fixed || (split > 1 && vmax == 0.0f)) return;
    __syncthreads();

    // 4. write-back (as brick_unit)
    const double inv = ldexp(1.0, -F);
    for (int i = tid; i < nvox; i += kFiledThreads) {
        const int k = i % ez, j = (i / ez) % ey, ii = i / (ez * ey);
        const int64_t gi = ((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k);
        const float v = (float)((double)(long long)acc[ii * kStrideX + j * kStrideY + k] * inv);
        if (split > 1)
            add_f32_ieee(o + gi, v);
        else
            o[gi] = accumulate ? o[gi] + v : v;
    }
}

template <int MODE>
__global__ __launch_bounds__(kFiledThreads) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_back_filed(
    const float* __restrict__ sino, float* __restrict__ out, int64_t n_ray, int64_t vol_elems, Dims3 dims,
    BrickGrid bg, int n_view, const int32_t* __restrict__ plan, const uint32_t* __restrict__ totals,
    const float* __restrict__ view_e, const float* __restrict__ src, const float* __restrict__ dst,
    const float* __restrict__ minv, const float* __restrict__ b, const float* __restrict__ view_max,
    const int64_t* __restrict__ offsets, int unit_slots, const uint32_t* __restrict__ ids, int accumulate,
    int split, const int32_t* __restrict__ order) {
    __shared__ unsigned long long acc[kBrickLds];
    __shared__ uint2 wq[kFiledThreads / 64 * kFileWaveQueue];
    __shared__ float s_red[kFiledThreads / 64];
    __shared__ float s_red2[kFiledThreads / 64];
    __shared__ uint32_t s_next;
    const int tid = threadIdx.x;
    const uint32_t ux = order ? (uint32_t)order[blockIdx.x] : blockIdx.x;
    if (ux >= gridDim.x) return;  // not a unit (order is meant to be a permutation of them)
    const int bid = (int)(ux / (uint32_t)split), part = (int)(ux % (uint32_t)split);
    const int64_t bi = blockIdx.y;
    const Window<3> win = brick_window(dims, bg, bid);
    for (int i = tid; i < kBrickLds; i += kFiledThreads) acc[i] = 0ull;
    if (tid == 0) s_next = 0u;

    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const float* __restrict__ sn = sino + bi * n_ray;
    const int32_t* __restrict__ p = plan + (int64_t)bid * n_view * kPlanInts;
    const uint32_t total = totals[bid];
    const int64_t* __restrict__ uo = offsets + (int64_t)bid * kFileHalves * unit_slots;
    const int64_t lo = (split == 1) ? uo[0] : uo[part * unit_slots];
    const int64_t hi = (split == 1) ? uo[kFileHalves * unit_slots] : uo[(part + 1) * unit_slots];
    float* __restrict__ o = out + bi * vol_elems;

    // 1. value bound over the brick's views with a non-empty rectangle and the
    // largest midpoint error bound over all views (as brick_unit)
    float vmax = 0.0f, emax = 0.0f;
    bool nonfinite = false;
    {
        const float* __restrict__ vm = view_max + bi * n_view;
        for (int v = tid; v < n_view; v += kFiledThreads) {
            if (p[(int64_t)v * kPlanInts + 2] > 0) {
                const float m = vm[v];
                nonfinite |= !(m <= 3.0e38f);
                vmax = fmaxf(vmax, m);
            }
            const float ev = view_e[v];
            emax = fmaxf(emax, (ev >= 0.0f && ev <= 3.0e38f) ? ev : __builtin_huge_valf());
        }
    }
    vmax = nonfinite ? __builtin_huge_valf() : vmax;
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
        vmax = fmaxf(vmax, __shfl_xor(vmax, o2));
        emax = fmaxf(emax, __shfl_xor(emax, o2));
    }
    if (tid % 64 == 0) {
        s_red[tid / 64] = vmax;
        s_red2[tid / 64] = emax;
    }
    __syncthreads();
    vmax = 0.0f;
    emax = 0.0f;
    for (int k = 0; k < kFiledThreads / 64; ++k) {
        vmax = fmaxf(vmax, s_red[k]);
        emax = fmaxf(emax, s_red2[k]);
    }
    vmax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vmax)));
    emax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(emax)));
    const float e_brick = emax * 1.0009765625f;

    filed_unit_rest<MODE>(g, win, dims, o, sn, src, dst, ids + lo, (uint32_t)(hi - lo), n_ray, vmax, e_brick,
                          total, accumulate, split, acc, &s_next, wq + (tid / 64) * kFileWaveQueue);
}

}  // namespace ctlab

namespace ctlab {

// Checks shared by the filed-list entry points; fills the plan's sections.
struct PlanParts {
    BrickGrid bg;
    int64_t nb, V;
    const uint32_t* totals;
    const float* view_e;
    const float* frames;
};

static int filed_check(const char* fn, int64_t n_view, int64_t det_rows, int64_t det_cols, const int32_t* plan,
                       const float* M_inv, const float* b, const int64_t* n, PlanParts* pp) {
    int st = check_grid3(fn, n);
    if (st) return st;
    CTLAB_REQUIRE(n_view >= 1 && det_rows >= 1 && det_cols >= 1 &&
                      n_view * det_rows * det_cols <= 0xffffffffLL - 4096,
                  "%s: bad detector layout (n_view*rows*cols must be below 2^32 - 4096)", fn);
    CTLAB_REQUIRE(plan && M_inv && b, "%s: null pointer", fn);
    pp->bg = brick_grid(n);
    pp->nb = (int64_t)pp->bg.nbx * pp->bg.nby * pp->bg.nbz;
    pp->V = n[0] * n[1] * n[2];
    pp->totals = reinterpret_cast<const uint32_t*>(plan + pp->nb * n_view * kPlanInts);
    pp->view_e = reinterpret_cast<const float*>(plan + pp->nb * (n_view * kPlanInts + 1));
    pp->frames = pp->view_e + n_view;
    return CTLAB_OK;
}

}  // namespace ctlab

extern "C" int64_t ctlab_brick_filed_slots_3d(int64_t n_x, int64_t n_y, int64_t n_z, int64_t n_view) {
    using namespace ctlab;
    if (n_x < 1 || n_y < 1 || n_z < 1 || n_view < 1) return -1;
    const int64_t n[3] = {n_x, n_y, n_z};
    const BrickGrid bg = brick_grid(n);
    return (int64_t)bg.nbx * bg.nby * bg.nbz * kFileHalves * filed_unit_slots(n_view);
}


extern "C" int ctlab_brick_filed_count_3d(const float* src, const float* dst, int64_t n_view, int64_t det_rows,
                                          int64_t det_cols, const int32_t* plan, const float* M_inv, const float* b,
                                          int64_t n_x, int64_t n_y, int64_t n_z, uint32_t* counts,
                                          ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "brick_filed_count_3d";
    const int64_t n[3] = {n_x, n_y, n_z};
    PlanParts pp;
    int st = filed_check(fn, n_view, det_rows, det_cols, plan, M_inv, b, n, &pp);
    if (st) return st;
    CTLAB_REQUIRE(src && dst && counts, "%s: null pointer", fn);
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    hipLaunchKernelGGL(k_file_units<false>, dim3((unsigned)(pp.nb * kFileHalves)), dim3(kFileThreads), 0,
                       as_stream(stream), src, dst, d, pp.bg, (int)n_view, (int)det_rows, (int)det_cols, plan,
                       pp.totals, pp.view_e, pp.frames, M_inv, b, filed_chunk_views(n_view),
                       filed_unit_slots(n_view), counts, nullptr, nullptr, (int64_t)0);
    return check_launch(fn);
}

extern "C" int ctlab_brick_filed_fill_3d(const float* src, const float* dst, int64_t n_view, int64_t det_rows,
                                         int64_t det_cols, const int32_t* plan, const float* M_inv, const float* b,
                                         int64_t n_x, int64_t n_y, int64_t n_z, const int64_t* offsets,
                                         uint32_t* ids, int64_t n_ids, ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "brick_filed_fill_3d";
    const int64_t n[3] = {n_x, n_y, n_z};
    PlanParts pp;
    int st = filed_check(fn, n_view, det_rows, det_cols, plan, M_inv, b, n, &pp);
    if (st) return st;
    CTLAB_REQUIRE(src && dst && offsets && ids && n_ids >= 0, "%s: null pointer", fn);
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    hipLaunchKernelGGL(k_file_units<true>, dim3((unsigned)(pp.nb * kFileHalves)), dim3(kFileThreads), 0,
                       as_stream(stream), src, dst, d, pp.bg, (int)n_view, (int)det_rows, (int)det_cols, plan,
                       pp.totals, pp.view_e, pp.frames, M_inv, b, filed_chunk_views(n_view),
                       filed_unit_slots(n_view), nullptr, offsets, ids, n_ids);
    return check_launch(fn);
}

extern "C" int ctlab_back_project_3d_brick_filed(const float* sino, int64_t batch, const float* src,
                                                 const float* dst, int64_t n_view, int64_t det_rows, int64_t det_cols,
                                                 const int32_t* plan, const int64_t* offsets, const uint32_t* ids,
                                                 const int32_t* order, const float* M_inv, const float* b,
                                                 int64_t n_x, int64_t n_y,
                                                 int64_t n_z, int mode, int accumulate, float* out, float* view_max,
                                                 ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "back_project_3d_brick_filed";
    const int64_t n[3] = {n_x, n_y, n_z};
    PlanParts pp;
    int st = filed_check(fn, n_view, det_rows, det_cols, plan, M_inv, b, n, &pp);
    if (st) return st;
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range", fn, (long long)batch);
    CTLAB_REQUIRE(mode == CTLAB_MODE_TORCH || mode == CTLAB_MODE_OTF, "%s: unknown mode %d", fn, mode);
    CTLAB_REQUIRE(out && sino && src && dst && offsets && ids && view_max, "%s: null pointer", fn);
    const int64_t n_ray = n_view * det_rows * det_cols;
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    const int split = accumulate ? 1 : kFileHalves;
    hipLaunchKernelGGL(k_view_absmax, dim3((unsigned)n_view, (unsigned)batch), dim3(256), 0, as_stream(stream), sino,
                       det_rows * det_cols, (int)n_view, view_max);
    st = check_launch(fn);
    if (st) return st;
    if (split > 1) {
        st = fill_f32(out, batch * pp.V, 0.0f, as_stream(stream));
        if (st) return st;
    }
    dim3 grid((unsigned)(pp.nb * split), (unsigned)batch);
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL(k_back_filed<MODE_TORCH>, grid, dim3(kFiledThreads), 0, as_stream(stream), sino, out, n_ray,
                           pp.V, d, pp.bg, (int)n_view, plan, pp.totals, pp.view_e, src, dst, M_inv, b, view_max,
                           offsets, filed_unit_slots(n_view), ids, accumulate, split, order);
    else
        hipLaunchKernelGGL(k_back_filed<MODE_OTF>, grid, dim3(kFiledThreads), 0, as_stream(stream), sino, out, n_ray,
                           pp.V, d, pp.bg, (int)n_view, plan, pp.totals, pp.view_e, src, dst, M_inv, b, view_max,
                           offsets, filed_unit_slots(n_view), ids, accumulate, split, order);
    return check_launch(fn);
}

// ===========================================================================
// Brick lists filed for the lean back projector
// ===========================================================================
//
// The brick lists above (rays in any order) are refiled once per ray set the
// way the plan's lists are: a brick's list becomes two units, the rays of even
// and of odd id (about the same work each), each cut into chunks of
// consecutive ray ids (ray sets in spatial order keep neighbouring rays
// together) and each chunk grouped by the 64 length classes of the rays' slab
// estimates, longest first.  k_back_filed_bins then traces them like
// k_back_filed.

namespace ctlab {

constexpr int kBinChunks = 64;  // id chunks per unit
constexpr int kBinBuckets = 64;  // eighth-octave length classes of the brick bins

static int64_t bin_chunk_rays(int64_t n_ray) {
    const int64_t half = (n_ray + 1) / 2;
    return std::max<int64_t>(1, (half + kBinChunks - 1) / kBinChunks);
}

static int bin_unit_slots(int64_t n_ray) {
    const int64_t half = (n_ray + 1) / 2, R = bin_chunk_rays(n_ray);
    return (int)std::max<int64_t>(1, (half + R - 1) / R) * kBinBuckets;
}

// Exclusive scan of v[0, n) in LDS (n <= 16 * 256) by a 256-thread block;
// returns the total.  Every thread of the block calls it.
__device__ uint32_t block_scan_256(uint32_t* v, int n, uint32_t* wsum) {
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    const int per = (n + 255) / 256;
    const int i0 = min(n, tid * per), i1 = min(n, i0 + per);
    uint32_t sum = 0;
    for (int i = i0; i < i1; ++i) sum += v[i];
    uint32_t x = sum;  // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t run = x - sum;
    for (int k = 0; k < w; ++k) run += wsum[k];
    const uint32_t total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    for (int i = i0; i < i1; ++i) {
        const uint32_t c = v[i];
        v[i] = run;
        run += c;
    }
    __syncthreads();
    return total;
}

// One block per brick: the brick's list ids[offsets[b] .. offsets[b+1]) is
// written to filed[] in the same range, slot by slot, with slot_offsets[b * 2S
// + k] the start of slot k (unit (b, ray & 1), chunk (ray >> 1) / chunk_rays,
// class); ids of n_ray or more (never filled) go to the brick's last slot.
__global__ __launch_bounds__(256) void k_bin_file(const float* __restrict__ src, const float* __restrict__ dst,
                                                  int64_t n_ray, Dims3 dims, BrickGrid bg,
                                                  const float* __restrict__ minv, const float* __restrict__ b,
                                                  const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ ids, const float* __restrict__ emax_in,
                                                  int64_t chunk_rays, int unit_slots,
                                                  int64_t* __restrict__ slot_offsets, uint32_t* __restrict__ filed) {
    __shared__ uint32_t hist[2 * kBinChunks * kBinBuckets];
    __shared__ uint32_t wsum[4];
    const int tid = threadIdx.x;
    const int bk = blockIdx.x, nslot = 2 * unit_slots;
    const uint32_t base = offsets[bk], cnt = offsets[bk + 1] - base;
    const Window<3> win = brick_window(dims, bg, bk);
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const float emax = emax_in[0];
    const float e_brick = (emax >= 0.0f && emax <= 3.0e38f) ? emax * 1.0009765625f : __builtin_huge_valf();
    const float m = ((e_brick <= 8.0f) ? e_brick : 8.0f) + 0.05f;
    float blo[3], bhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        blo[a] = (float)win.lo[a] - 0.5f - m;
        bhi[a] = (float)(win.lo[a] + win.ext[a]) - 0.5f + m;
    }
    auto key = [&](uint32_t ray) -> int {
        if ((int64_t)ray >= n_ray) return nslot - 1;
        const uint32_t est = segment_estimate<MODE_TORCH>(g, src + 3ull * ray, dst + 3ull * ray, blo, bhi);
        uint32_t bc = 0;
        if (est) {
            const uint32_t l2 = 31u - (uint32_t)__clz(est);
            bc = min(8u * l2 + ((est << 3) >> l2) - 8u, (uint32_t)(kBinBuckets - 1));
        }
        const int ch = (int)((int64_t)(ray >> 1) / chunk_rays);
        return (int)(ray & 1u) * unit_slots + ch * kBinBuckets + ((ch & 1) ? (int)bc : kBinBuckets - 1 - (int)bc);
    };
    for (int i = tid; i < nslot; i += 256) hist[i] = 0u;
    __syncthreads();
    for (uint32_t i = tid; i < cnt; i += 256) atomicAdd(&hist[key(ids[base + i])], 1u);
    __syncthreads();
    block_scan_256(hist, nslot, wsum);
    for (int i = tid; i < nslot; i += 256) {
        hist[i] += base;
        slot_offsets[(int64_t)bk * nslot + i] = hist[i];
    }
    if (bk == (int)gridDim.x - 1 && tid == 0) slot_offsets[(int64_t)gridDim.x * nslot] = base + cnt;
    __syncthreads();
    for (uint32_t i = tid; i < cnt; i += 256) {
        const uint32_t ray = ids[base + i];
        filed[atomicAdd(&hist[key(ray)], 1u)] = ray;
    }
}

// One workgroup per unit (brick, ray parity; with split 1 both parities, one
// contiguous range).  Value bound: the batch element's largest |value|; ray
// count bound: the brick's list length; midpoint error bound: the largest over
// the ray set.
template <int MODE>
__global__ __launch_bounds__(kFiledThreads) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_back_filed_bins(
    const float* __restrict__ sino, float* __restrict__ out, int64_t n_ray, int64_t vol_elems, Dims3 dims,
    BrickGrid bg, const int64_t* __restrict__ slot_offsets, int unit_slots, const uint32_t* __restrict__ filed,
    const float* __restrict__ emax_in, const float* __restrict__ src, const float* __restrict__ dst,
    const float* __restrict__ minv, const float* __restrict__ b, const float* __restrict__ vmax_in,
    int accumulate, int split, const int32_t* __restrict__ order) {
    __shared__ unsigned long long acc[kBrickLds];
    __shared__ uint2 wq[kFiledThreads / 64 * kFileWaveQueue];
    __shared__ uint32_t s_next;
    const int tid = threadIdx.x;
    const uint32_t ux = order ? (uint32_t)order[blockIdx.x] : blockIdx.x;
    if (ux >= gridDim.x) return;  // not a unit (order is meant to be a permutation of them)
    const int bid = (int)(ux / (uint32_t)split), part = (int)(ux % (uint32_t)split);
    const int64_t bi = blockIdx.y;
    const Window<3> win = brick_window(dims, bg, bid);
    for (int i = tid; i < kBrickLds; i += kFiledThreads) acc[i] = 0ull;
    if (tid == 0) s_next = 0u;
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const int64_t* __restrict__ uo = slot_offsets + (int64_t)bid * 2 * unit_slots;
    const int64_t lo = (split == 1) ? uo[0] : uo[part * unit_slots];
    const int64_t hi = (split == 1) ? uo[2 * unit_slots] : uo[(part + 1) * unit_slots];
    const uint32_t total = (uint32_t)(uo[2 * unit_slots] - uo[0]);
    const float vmax = (total > 0) ? vmax_in[bi] : 0.0f;
    const float emax = emax_in[0];
    const float e_brick = (emax >= 0.0f && emax <= 3.0e38f) ? emax * 1.0009765625f : __builtin_huge_valf();
    __syncthreads();
    filed_unit_rest<MODE>(g, win, dims, out + bi * vol_elems, sino + bi * n_ray, src, dst, filed + lo,
                          (uint32_t)(hi - lo), n_ray, vmax, e_brick, total, accumulate, split, acc, &s_next,
                          wq + (tid / 64) * kFileWaveQueue);
}

}  // namespace ctlab

extern "C" int64_t ctlab_brick_bins_slots_3d(int64_t n_x, int64_t n_y, int64_t n_z, int64_t n_ray) {
    using namespace ctlab;
    if (n_x < 1 || n_y < 1 || n_z < 1 || n_ray < 0) return -1;
    const int64_t n[3] = {n_x, n_y, n_z};
    const BrickGrid bg = brick_grid(n);
    return (int64_t)bg.nbx * bg.nby * bg.nbz * 2 * bin_unit_slots(n_ray);
}

extern "C" int ctlab_brick_bins_file_3d(const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                        const float* b, int64_t n_x, int64_t n_y, int64_t n_z,
                                        const uint32_t* offsets, const uint32_t* ids, const float* emax,
                                        int64_t* slot_offsets, uint32_t* filed, ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "brick_bins_file_3d";
    const int64_t n[3] = {n_x, n_y, n_z};
    int st = check_grid3(fn, n);
    if (st) return st;
    CTLAB_REQUIRE(n_ray >= 0 && n_ray <= 0xffffffffLL, "%s: n_ray %lld out of range", fn, (long long)n_ray);
    CTLAB_REQUIRE(offsets && ids && emax && slot_offsets && filed && M_inv && b && (n_ray == 0 || (src && dst)),
                  "%s: null pointer", fn);
    const BrickGrid bg = brick_grid(n);
    const int64_t nbk = (int64_t)bg.nbx * bg.nby * bg.nbz;
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    hipLaunchKernelGGL(k_bin_file, dim3((unsigned)nbk), dim3(256), 0, as_stream(stream), src, dst, n_ray, d, bg, M_inv,
                       b, offsets, ids, emax, bin_chunk_rays(n_ray), bin_unit_slots(n_ray), slot_offsets, filed);
    return check_launch(fn);
}

extern "C" int ctlab_back_project_3d_brick_bins(const float* sino, int64_t batch, const float* src, const float* dst,
                                                int64_t n_ray, const int64_t* slot_offsets, const uint32_t* filed,
                                                const float* emax, const int32_t* order, const float* M_inv,
                                                const float* b, int64_t n_x, int64_t n_y, int64_t n_z, int mode,
                                                int accumulate, float* out, float* vmax, ctlab_stream_t stream) {
    using namespace ctlab;
    const char* fn = "back_project_3d_brick_bins";
    const int64_t n[3] = {n_x, n_y, n_z};
    int st = check_grid3(fn, n);
    if (st) return st;
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range", fn, (long long)batch);
    CTLAB_REQUIRE(mode == CTLAB_MODE_TORCH || mode == CTLAB_MODE_OTF, "%s: unknown mode %d", fn, mode);
    CTLAB_REQUIRE(n_ray >= 0 && n_ray <= 0xffffffffLL, "%s: n_ray %lld out of range", fn, (long long)n_ray);
    CTLAB_REQUIRE(out && slot_offsets && filed && emax && vmax && M_inv && b, "%s: null pointer", fn);
    CTLAB_REQUIRE(n_ray == 0 || (sino && src && dst), "%s: null pointer", fn);
    const BrickGrid bg = brick_grid(n);
    const int64_t nbk = (int64_t)bg.nbx * bg.nby * bg.nbz;
    const int64_t V = n_x * n_y * n_z;
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    st = fill_f32(vmax, batch, 0.0f, as_stream(stream));
    if (st) return st;
    if (n_ray > 0) {
        const unsigned blocks = (unsigned)std::min<int64_t>((n_ray + 255) / 256, 1024);
        hipLaunchKernelGGL(k_absmax_all, dim3(blocks, (unsigned)batch), dim3(256), 0, as_stream(stream), sino, n_ray,
                           reinterpret_cast<unsigned int*>(vmax));
        st = check_launch(fn);
        if (st) return st;
    }
    const int split = accumulate ? 1 : 2;
    if (split > 1) {
        st = fill_f32(out, batch * V, 0.0f, as_stream(stream));
        if (st) return st;
    }
    dim3 grid((unsigned)(nbk * split), (unsigned)batch);
    const int slots = bin_unit_slots(n_ray);
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL(k_back_filed_bins<MODE_TORCH>, grid, dim3(kFiledThreads), 0, as_stream(stream), sino, out,
                           n_ray, V, d, bg, slot_offsets, slots, filed, emax, src, dst, M_inv, b, vmax, accumulate,
                           split, order);
    else
        hipLaunchKernelGGL(k_back_filed_bins<MODE_OTF>, grid, dim3(kFiledThreads), 0, as_stream(stream), sino, out,
                           n_ray, V, d, bg, slot_offsets, slots, filed, emax, src, dst, M_inv, b, vmax, accumulate,
                           split, order);
    return check_launch(fn);
}
