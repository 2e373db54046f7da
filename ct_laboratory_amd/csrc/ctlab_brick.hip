// ctlab_brick.hip -- atomic-free 3D back projection for rays that come as a
// stack of flat-panel views ([view][rows][cols], one source per view), the
// layout of the cone-beam and static-ring builders.
//
// A workgroup owns a 32^3 brick of the output volume, accumulated in LDS
// (128 KB), and loops over every view.  Per view it projects the brick's
// corners from the view's source onto the detector plane and takes the padded
// bounding rectangle of detector pixels (the brick's shadow); every ray whose
// rounded voxels touch the brick lies in that rectangle.  Each such ray is
// traced with the exact traversal clipped to the brick window
// (trace_ray_window): the brick receives exactly the segments whose voxel it
// contains, added with LDS float atomics (ds_add_f32).  The brick is written
// once with plain stores: no global atomics at all.
//
// Exactness does not depend on the detector model: the frame of a view is only
// used to bound the shadow, conservatively.  A view whose rays do not fit the
// flat-panel model within a quarter pixel (source moving, detector not
// affine) is marked invalid and uses its whole detector for every brick.
#include "ctlab_common.h"
#include "siddon_traverse.h"

namespace ctlab {

constexpr int kBrick = 32;                  // brick edge, voxels
constexpr int kBrickThreads = 1024;         // 16 waves
constexpr int kViewChunk = kBrickThreads;   // views whose shadows sit in LDS at once
constexpr int kFrameFloats = 16;
// LDS strides of the brick: padded so that 32 lanes stepping along any one
// axis hit 32 different banks (33 and 32*33+1 are odd multiples mod 32).
constexpr int kStrideZ = 1, kStrideY = kBrick + 1, kStrideX = kBrick * (kBrick + 1) + 1;
constexpr int kBrickLds = kBrick * kStrideX;

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
    // Any ray with a segment whose rounded voxel is in the box passes within E
    // of it; the model line S -> O + rR + cC of its pixel deviates from the ray
    // by at most the source and pixel deviations.  Margin: one voxel more.
    // Float precision is ample: coordinates of ~1e3 voxels err by ~1e-4 voxels,
    // far inside the margin and the one-pixel pad.
    const float m = 1.0f + f[F_E] + f[F_DEVS] + f[F_DEVD];
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
    const float pad = 1.0f;
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

template <int MODE>
__global__ __launch_bounds__(kBrickThreads) void k_back_brick(
    const float* __restrict__ sino, float* __restrict__ out, int64_t n_ray, int64_t vol_elems, Dims3 dims,
    int nbx, int nby, int nbz, int n_view, int rows, int cols, int view_chunk, const float* __restrict__ frames,
    const float* __restrict__ src, const float* __restrict__ dst, const float* __restrict__ minv,
    const float* __restrict__ b, int accumulate) {
    __shared__ float acc[kBrickLds];
    __shared__ int s_r0[kViewChunk], s_w[kViewChunk], s_c0[kViewChunk];
    __shared__ uint32_t s_pre[kViewChunk + 1];
    __shared__ uint32_t s_wsum[kBrickThreads / 64];

    const int tid = threadIdx.x;
    const int bid = blockIdx.x;
    const int bz = bid % nbz, by = (bid / nbz) % nby, bx = bid / (nbz * nby);
    Window<3> win;
    win.lo[0] = bx * kBrick;
    win.lo[1] = by * kBrick;
    win.lo[2] = bz * kBrick;
    for (int a = 0; a < 3; ++a) win.ext[a] = min(kBrick, dims.n[a] - win.lo[a]);
    win.stride[0] = kStrideX;
    win.stride[1] = kStrideY;
    win.stride[2] = kStrideZ;
    for (int i = tid; i < kBrickLds; i += kBrickThreads) acc[i] = 0.0f;

    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const int64_t bi = blockIdx.y;
    const float* __restrict__ sn = sino + bi * n_ray;

    for (int v0 = 0; v0 < n_view; v0 += view_chunk) {
        const int nv = min(view_chunk, n_view - v0);
        // 1. shadow rectangle of this brick in each view of the chunk
        uint32_t cnt = 0;
        if (tid < nv) {
            int r0, r1, c0, c1;
            if (!shadow_rect(frames + (int64_t)(v0 + tid) * kFrameFloats, win.lo, win.ext, rows, cols, &r0, &r1,
                             &c0, &c1)) {
                r0 = 0;
                r1 = rows - 1;
                c0 = 0;
                c1 = cols - 1;
            }
            const int h = r1 - r0 + 1, w = c1 - c0 + 1;
            s_r0[tid] = r0;
            s_w[tid] = w > 0 ? w : 0;
            s_c0[tid] = c0;
            cnt = (h > 0 && w > 0) ? (uint32_t)h * (uint32_t)w : 0u;
        }
        // 2. exclusive prefix sum of the counts over the block
        uint32_t x = cnt;
        const int lane = tid % 64, wv = tid / 64;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_wsum[wv] = x;
        __syncthreads();
        if (tid < 64) {
            uint32_t ws = (lane < kBrickThreads / 64) ? s_wsum[lane] : 0u;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(ws, o);
                if (lane >= o) ws += y;
            }
            if (lane < kBrickThreads / 64) s_wsum[lane] = ws;  // inclusive
        }
        __syncthreads();
        const uint32_t incl = x + (wv > 0 ? s_wsum[wv - 1] : 0u);
        if (tid < nv) s_pre[tid + 1] = incl;
        if (tid == 0) s_pre[0] = 0;
        __syncthreads();
        const uint32_t total = s_pre[nv];

        // 3. trace every (view, pixel) of the chunk's shadows, columns fastest
        //    (neighbouring lanes read neighbouring src/dst/sino entries)
        for (uint64_t item = tid; item < total; item += kBrickThreads) {
            int lo = 0, hi = nv;  // largest view with s_pre[view] <= item
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pre[mid] <= item) lo = mid; else hi = mid;
            }
            const uint32_t local = (uint32_t)(item - s_pre[lo]);
            const uint32_t w = (uint32_t)s_w[lo];
            const int r = s_r0[lo] + (int)(local / w);
            const int c = s_c0[lo] + (int)(local % w);
            const int64_t ray = ((int64_t)(v0 + lo) * rows + r) * cols + c;
            const float val = sn[ray];
            if (val == 0.0f) continue;
            float s[3], d[3];
            for (int a = 0; a < 3; ++a) {
                s[a] = src[3 * ray + a];
                d[a] = dst[3 * ray + a];
            }
            auto add = [&](uint32_t li, float len) { atomicAdd(&acc[li], val * len); };
            if (g.diag)
                trace_ray_window<3, MODE, true>(g, win, s, d, add);
            else
                trace_ray_window<3, MODE, false>(g, win, s, d, add);
        }
        __syncthreads();
    }

    // 4. write the brick (k fastest: coalesced rows of the volume)
    float* __restrict__ o = out + bi * vol_elems;
    const int ey = win.ext[1], ez = win.ext[2];
    const int nvox = win.ext[0] * ey * ez;
    for (int i = tid; i < nvox; i += kBrickThreads) {
        const int k = i % ez, j = (i / ez) % ey, ii = i / (ez * ey);
        const int64_t gi = ((int64_t)(win.lo[0] + ii) * dims.n[1] + (win.lo[1] + j)) * dims.n[2] + (win.lo[2] + k);
        const float v = acc[ii * kStrideX + j * kStrideY + k];
        o[gi] = accumulate ? o[gi] + v : v;
    }
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

extern "C" int ctlab_back_project_3d_brick(const float* sino, int64_t batch, const float* src, const float* dst,
                                           int64_t n_view, int64_t det_rows, int64_t det_cols, const float* frames,
                                           const float* M_inv, const float* b, int64_t n_x, int64_t n_y,
                                           int64_t n_z, int mode, int accumulate, float* out,
                                           ctlab_stream_t stream) {
    const char* fn = "back_project_3d_brick";
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
    CTLAB_REQUIRE(n_view >= 0 && det_rows >= 1 && det_cols >= 1 && det_rows * det_cols <= 0x7fffffffLL,
                  "%s: bad detector layout", fn);
    CTLAB_REQUIRE(out, "%s: out is null", fn);
    const int64_t n_ray = n_view * det_rows * det_cols;
    if (n_ray > 0) CTLAB_REQUIRE(sino && src && dst && frames && M_inv && b, "%s: null pointer", fn);
    // views per chunk: the chunk's pixel count must fit in 32 bits
    int64_t chunk = kViewChunk;
    while (chunk > 1 && chunk * det_rows * det_cols > 0xffffffffLL) chunk /= 2;
    const int64_t nb = ((n_x + kBrick - 1) / kBrick) * ((n_y + kBrick - 1) / kBrick) * ((n_z + kBrick - 1) / kBrick);
    Dims3 d = {{(int)n_x, (int)n_y, (int)n_z}};
    dim3 grid((unsigned)nb, (unsigned)batch);
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL(k_back_brick<MODE_TORCH>, grid, dim3(kBrickThreads), 0, as_stream(stream), sino, out,
                           n_ray, V, d, (int)((n_x + kBrick - 1) / kBrick), (int)((n_y + kBrick - 1) / kBrick),
                           (int)((n_z + kBrick - 1) / kBrick), (int)n_view, (int)det_rows, (int)det_cols, (int)chunk,
                           frames, src, dst, M_inv, b, accumulate);
    else
        hipLaunchKernelGGL(k_back_brick<MODE_OTF>, grid, dim3(kBrickThreads), 0, as_stream(stream), sino, out,
                           n_ray, V, d, (int)((n_x + kBrick - 1) / kBrick), (int)((n_y + kBrick - 1) / kBrick),
                           (int)((n_z + kBrick - 1) / kBrick), (int)n_view, (int)det_rows, (int)det_cols, (int)chunk,
                           frames, src, dst, M_inv, b, accumulate);
    return check_launch(fn);
}
