// ctlab_otf.hip -- on-the-fly Siddon forward/back projection (2D and 3D) and
// the per-ray segment-list instrument.
//
// One lane traces one ray (siddon_traverse.h).  The forward kernel gathers
// volume values along the ray and writes one line integral per lane; the back
// kernel scatters sino*length with float atomic adds (global_atomic_add_f32,
// no return value).  2D rays may also take the reference's 2D on-the-fly march
// (MODE_MARCH, march_ray_2d), whose back projection always uses the atomics.  In the forward, blockIdx.y is a group of up to NB batch
// members: a lane traces its ray once and gathers from each member's volume,
// so the traversal arithmetic (which bounds the kernel) is shared by the
// group.  In the back projection blockIdx.y is the batch index.
#include "ctlab_common.h"
#include "siddon_traverse.h"

namespace ctlab {

template <int ND>
struct Dims {
    int n[ND];
};

// Lane -> ray schedule.  With a detector layout (rays ordered
// [view][row][col], col fastest), a block is a tile of 64 rows x 4 columns of
// one view and each wave takes 16 consecutive rows of the 4 columns.  Rays of
// one column share their in-plane path and differ in the slow z direction,
// which is the fastest-varying volume axis, so one wave's gathers or atomics
// touch a few cache lines per column instead of 64.  Blocks are remapped so that each XCD (the
// hardware deals blocks round-robin to the 8 XCDs) receives a contiguous run
// of tiles and neighbouring tiles share that XCD's L2.
struct RaySched {
    int64_t n_ray;
    uint32_t rows, cols;        // 0: plain order, block b covers rays [256 b, 256 b + 256)
    uint32_t tiles_r, tiles_c;  // ceil(rows / 64), ceil(cols / 4)
    uint32_t n_view;            // n_ray / (rows * cols)
    uint32_t n_blocks;          // gridDim.x
};

constexpr int kTileRows = 64;
constexpr int kTileCols = kBlock / kTileRows;
// A wave covers kWaveRows rows x (64 / kWaveRows) columns of its block's tile:
// 16 x 4 keeps the wave's rays close in length (92% of the lanes busy in the
// segment loop on the 256^3 cone-beam workload, against 79% for 64 x 1) while
// its gathers still touch few cache lines (16 z-adjacent rays per column).
constexpr int kWaveRows = 16;
constexpr uint32_t kXcds = 8;
// Consecutive blocks of one XCD take the same detector tile in kViewGroup
// consecutive views: those rays cross nearly the same voxels, so the volume
// lines one of them brings into the XCD's L2 serve the others.
constexpr uint32_t kViewGroup = 8;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nb) {
    const uint32_t q = nb / kXcds, r = nb % kXcds, x = bid % kXcds, j = bid / kXcds;
    return (x < r) ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

__device__ __forceinline__ int64_t ray_of_lane(const RaySched& s) {
    const uint32_t t = xcd_remap(blockIdx.x, s.n_blocks);
    if (s.rows == 0) {
        const int64_t r = (int64_t)t * kBlock + threadIdx.x;
        return r < s.n_ray ? r : -1;
    }
    // t -> (group of kViewGroup views, tile, view within the group); the last
    // group may hold fewer views.
    const uint32_t per_view = s.tiles_r * s.tiles_c;
    const uint32_t grp = t / (per_view * kViewGroup);
    const uint32_t rem = t - grp * per_view * kViewGroup;
    const uint32_t gv = min(kViewGroup, s.n_view - grp * kViewGroup);
    const uint32_t tile = rem / gv;
    const uint32_t view = grp * kViewGroup + (rem - tile * gv);
    const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
    constexpr uint32_t kWavesDown = kTileRows / kWaveRows;  // waves stacked along the rows
    const uint32_t row = (tile / s.tiles_c) * kTileRows + (w % kWavesDown) * kWaveRows + l % kWaveRows;
    const uint32_t col = (tile % s.tiles_c) * kTileCols + (w / kWavesDown) * (64 / kWaveRows) + l / kWaveRows;
    if (row >= s.rows || col >= s.cols) return -1;
    return ((int64_t)view * s.rows + row) * s.cols + col;
}

template <int ND>
__device__ __forceinline__ void load_ray(const float* __restrict__ src, const float* __restrict__ dst, int64_t r,
                                         float* s, float* d) {
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        s[a] = src[r * ND + a];
        d[a] = dst[r * ND + a];
    }
}

template <int ND, int MODE, int NB>
__global__ __launch_bounds__(kBlock) void k_forward_otf(const float* __restrict__ vol, float* __restrict__ out,
                                                        RaySched sched, int64_t vol_elems, Dims<ND> dims,
                                                        const float* __restrict__ src,
                                                        const float* __restrict__ dst,
                                                        const float* __restrict__ minv,
                                                        const float* __restrict__ b, int zero_rule, int batch) {
    const int64_t r = ray_of_lane(sched);
    if (r < 0) return;
    const int b0 = (int)blockIdx.y * NB;
    const int64_t n_ray = sched.n_ray;
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, zero_rule);
    float s[ND], d[ND];
    load_ray<ND>(src, dst, r, s, d);
    // Members past the end of the batch (last group) read the last member's
    // volume and are not written.
    const float* __restrict__ v[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) v[k] = vol + (int64_t)min(b0 + k, batch - 1) * vol_elems;
    // The gather of a segment is consumed at the next visited segment, so its
    // memory latency overlaps the next crossing and midpoint computation.  Each
    // member's sum is formed in segment order (acc = fma(value, len, acc)),
    // whatever NB is.
    float acc[NB], pv[NB], pl = 0.0f;
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k] = pv[k] = 0.0f;
    auto gather = [&](uint32_t idx, float len) {
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            acc[k] = __builtin_fmaf(pv[k], pl, acc[k]);
            pv[k] = v[k][idx];
        }
        pl = len;
    };
    if constexpr (MODE == MODE_MARCH) {
        march_ray_2d(g, s, d, gather);
    } else if constexpr (MODE == MODE_TORCH) {
        if (g.pow2)
            trace_ray<ND, MODE, true, true>(g, s, d, gather);
        else if (g.diag)
            trace_ray<ND, MODE, true>(g, s, d, gather);
        else
            trace_ray<ND, MODE, false>(g, s, d, gather);
    } else {
        if (g.diag)
            trace_ray<ND, MODE, true>(g, s, d, gather);
        else
            trace_ray<ND, MODE, false>(g, s, d, gather);
    }
#pragma unroll
    for (int k = 0; k < NB; ++k)
        if (b0 + k < batch) out[(int64_t)(b0 + k) * n_ray + r] = __builtin_fmaf(pv[k], pl, acc[k]);
}

// Batch members per trace.  Four share a trace only while their four volumes
// stay well inside the 256 MB Infinity Cache (at 256^3, four volumes gathered
// together ran slower per volume than two); otherwise two.
constexpr int64_t kGroup4Bytes = 128ll << 20;

template <int ND, int MODE>
static void launch_forward_otf(const float* vol, float* out, const RaySched& sched, int64_t V, Dims<ND> dims,
                               const float* src, const float* dst, const float* minv, const float* b, int zero_rule,
                               int64_t batch, hipStream_t stream) {
    const int bt = (int)batch;
    if (batch >= 4 && 4 * V * (int64_t)sizeof(float) <= kGroup4Bytes)
        hipLaunchKernelGGL((k_forward_otf<ND, MODE, 4>), dim3(sched.n_blocks, (unsigned)((batch + 3) / 4)),
                           dim3(kBlock), 0, stream, vol, out, sched, V, dims, src, dst, minv, b, zero_rule, bt);
    else if (batch >= 2)
        hipLaunchKernelGGL((k_forward_otf<ND, MODE, 2>), dim3(sched.n_blocks, (unsigned)((batch + 1) / 2)),
                           dim3(kBlock), 0, stream, vol, out, sched, V, dims, src, dst, minv, b, zero_rule, bt);
    else
        hipLaunchKernelGGL((k_forward_otf<ND, MODE, 1>), dim3(sched.n_blocks, 1u), dim3(kBlock), 0, stream, vol,
                           out, sched, V, dims, src, dst, minv, b, zero_rule, bt);
}

template <int ND, int MODE>
__global__ __launch_bounds__(kBlock) void k_back_otf(const float* __restrict__ sino, float* __restrict__ out,
                                                     RaySched sched, int64_t vol_elems, Dims<ND> dims,
                                                     const float* __restrict__ src, const float* __restrict__ dst,
                                                     const float* __restrict__ minv, const float* __restrict__ b,
                                                     int zero_rule) {
    const int64_t r = ray_of_lane(sched);
    if (r < 0) return;
    const int64_t bi = blockIdx.y;
    const float val = sino[bi * sched.n_ray + r];
    if (val == 0.0f) return;  // every contribution would be exactly zero
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, zero_rule);
    float s[ND], d[ND];
    load_ray<ND>(src, dst, r, s, d);
    float* __restrict__ o = out + bi * vol_elems;
    auto scatter = [&](uint32_t idx, float len) { unsafeAtomicAdd(o + idx, val * len); };
    if constexpr (MODE == MODE_MARCH)
        march_ray_2d(g, s, d, scatter);
    else if (g.diag)
        trace_ray<ND, MODE, true>(g, s, d, scatter);
    else
        trace_ray<ND, MODE, false>(g, s, d, scatter);
}

template <int ND, int MODE>
__global__ __launch_bounds__(kBlock) void k_trace(int64_t n_ray, Dims<ND> dims, const float* __restrict__ src,
                                                  const float* __restrict__ dst, const float* __restrict__ minv,
                                                  const float* __restrict__ b, int zero_rule, int64_t cap,
                                                  int32_t* __restrict__ idx_out, float* __restrict__ len_out,
                                                  int32_t* __restrict__ counts) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ray) return;
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, zero_rule);
    float s[ND], d[ND];
    load_ray<ND>(src, dst, r, s, d);
    int32_t* io = idx_out + r * cap;
    float* lo = len_out + r * cap;
    int64_t k = 0;
    auto record = [&](uint32_t idx, float len) {
        if (k < cap) {
            io[k] = (int32_t)idx;
            lo[k] = len;
        }
        ++k;
    };
    if constexpr (MODE == MODE_MARCH)
        march_ray_2d(g, s, d, record);
    else if (g.diag)
        trace_ray<ND, MODE, true>(g, s, d, record);
    else
        trace_ray<ND, MODE, false>(g, s, d, record);
    counts[r] = (int32_t)k;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

template <int ND>
static int check_common(const char* fn, const int64_t* n, int64_t batch, int64_t n_ray, const void* src,
                        const void* dst, const void* minv, const void* b, int mode) {
    int64_t vox = 1;
    for (int a = 0; a < ND; ++a) {
        CTLAB_REQUIRE(n[a] >= 1 && n[a] <= 0x7fffffffLL, "%s: grid dimension %d is %lld (must be >= 1)", fn, a,
                      (long long)n[a]);
        vox *= n[a];
        CTLAB_REQUIRE(vox <= kMaxVoxels, "%s: grid has more than 2^31-1 voxels", fn);
    }
    int st = check_index_limits(fn, n, ND);
    if (st) return st;
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range [1, 65535]", fn, (long long)batch);
    CTLAB_REQUIRE(n_ray >= 0, "%s: n_ray %lld < 0", fn, (long long)n_ray);
    CTLAB_REQUIRE(grid_for(n_ray) <= kMaxGridX, "%s: too many rays", fn);
    CTLAB_REQUIRE(mode == CTLAB_MODE_TORCH || mode == CTLAB_MODE_OTF || (ND == 2 && mode == CTLAB_MODE_MARCH),
                  "%s: unknown mode %d", fn, mode);
    if (n_ray > 0) {
        CTLAB_REQUIRE(src && dst, "%s: src/dst is null", fn);
        CTLAB_REQUIRE(minv && b, "%s: M_inv/b is null", fn);
    }
    return CTLAB_OK;
}

template <int ND>
static Dims<ND> make_dims(const int64_t* n) {
    Dims<ND> d;
    for (int a = 0; a < ND; ++a) d.n[a] = (int)n[a];
    return d;
}

static int make_sched(const char* fn, int64_t n_ray, int64_t rows, int64_t cols, RaySched* s) {
    s->n_ray = n_ray;
    s->rows = s->cols = s->tiles_r = s->tiles_c = s->n_view = 0;
    int64_t blocks;
    if (rows <= 0 && cols <= 0) {
        blocks = (n_ray + kBlock - 1) / kBlock;
    } else {
        CTLAB_REQUIRE(rows > 0 && cols > 0 && rows <= 0x7fffffffLL && cols <= 0x7fffffffLL,
                      "%s: detector layout %lld x %lld invalid", fn, (long long)rows, (long long)cols);
        CTLAB_REQUIRE(n_ray % (rows * cols) == 0, "%s: n_ray %lld is not a multiple of the detector size %lld x %lld",
                      fn, (long long)n_ray, (long long)rows, (long long)cols);
        s->rows = (uint32_t)rows;
        s->cols = (uint32_t)cols;
        s->tiles_r = (uint32_t)((rows + kTileRows - 1) / kTileRows);
        s->tiles_c = (uint32_t)((cols + kTileCols - 1) / kTileCols);
        CTLAB_REQUIRE(n_ray / (rows * cols) <= 0xffffffffLL, "%s: too many views", fn);
        s->n_view = (uint32_t)(n_ray / (rows * cols));
        blocks = (n_ray / (rows * cols)) * (int64_t)s->tiles_r * s->tiles_c;
    }
    CTLAB_REQUIRE(blocks <= 0xffffffffLL && blocks <= kMaxGridX, "%s: too many rays for one launch", fn);
    s->n_blocks = (uint32_t)blocks;
    return CTLAB_OK;
}

template <int ND>
static int forward_otf(const char* fn, const float* vol, int64_t batch, const int64_t* n, const float* src,
                       const float* dst, int64_t n_ray, const float* minv, const float* b, int mode,
                       int zero_rule, int64_t rows, int64_t cols, float* out, hipStream_t stream) {
    int st = check_common<ND>(fn, n, batch, n_ray, src, dst, minv, b, mode);
    if (st) return st;
    RaySched sched;
    st = make_sched(fn, n_ray, rows, cols, &sched);
    if (st) return st;
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(vol && out, "%s: volume/out is null", fn);
    int64_t V = 1;
    for (int a = 0; a < ND; ++a) V *= n[a];
    if constexpr (ND == 2) {
        if (mode == CTLAB_MODE_MARCH) {
            launch_forward_otf<2, MODE_MARCH>(vol, out, sched, V, make_dims<2>(n), src, dst, minv, b, zero_rule,
                                              batch, stream);
            return check_launch(fn);
        }
    }
    if (mode == CTLAB_MODE_TORCH)
        launch_forward_otf<ND, MODE_TORCH>(vol, out, sched, V, make_dims<ND>(n), src, dst, minv, b, zero_rule, batch,
                                           stream);
    else
        launch_forward_otf<ND, MODE_OTF>(vol, out, sched, V, make_dims<ND>(n), src, dst, minv, b, zero_rule, batch,
                                         stream);
    return check_launch(fn);
}

template <int ND>
static int back_otf(const char* fn, const float* sino, int64_t batch, int64_t n_ray, const float* src,
                    const float* dst, const float* minv, const float* b, const int64_t* n, int mode,
                    int zero_rule, int64_t rows, int64_t cols, int accumulate, float* out, hipStream_t stream) {
    int st = check_common<ND>(fn, n, batch, n_ray, src, dst, minv, b, mode);
    if (st) return st;
    RaySched sched;
    st = make_sched(fn, n_ray, rows, cols, &sched);
    if (st) return st;
    CTLAB_REQUIRE(out, "%s: out is null", fn);
    if constexpr (ND == 2) {  // atomic-free: LDS tiles (ctlab_tile2d.hip); the march rule takes atomics
        if (n_ray > 0 && mode != CTLAB_MODE_MARCH) {
            CTLAB_REQUIRE(sino, "%s: sino is null", fn);
            return back_project_2d_tiles(fn, sino, batch, n_ray, src, dst, minv, b, n, mode, zero_rule, accumulate,
                                         out, stream);
        }
    }
    int64_t V = 1;
    for (int a = 0; a < ND; ++a) V *= n[a];
    if (!accumulate) {
        st = fill_f32(out, V * batch, 0.0f, stream);
        if (st) return st;
    }
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(sino, "%s: sino is null", fn);
    dim3 grid(sched.n_blocks, (unsigned)batch);
    if constexpr (ND == 2) {
        if (mode == CTLAB_MODE_MARCH) {
            hipLaunchKernelGGL((k_back_otf<2, MODE_MARCH>), grid, dim3(kBlock), 0, stream, sino, out, sched, V,
                               make_dims<2>(n), src, dst, minv, b, zero_rule);
            return check_launch(fn);
        }
    }
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL((k_back_otf<ND, MODE_TORCH>), grid, dim3(kBlock), 0, stream, sino, out, sched, V,
                           make_dims<ND>(n), src, dst, minv, b, zero_rule);
    else
        hipLaunchKernelGGL((k_back_otf<ND, MODE_OTF>), grid, dim3(kBlock), 0, stream, sino, out, sched, V,
                           make_dims<ND>(n), src, dst, minv, b, zero_rule);
    return check_launch(fn);
}

template <int ND>
static int trace(const char* fn, const int64_t* n, const float* src, const float* dst, int64_t n_ray,
                 const float* minv, const float* b, int mode, int zero_rule, int64_t cap, int32_t* idx,
                 float* len, int32_t* counts, hipStream_t stream) {
    int st = check_common<ND>(fn, n, 1, n_ray, src, dst, minv, b, mode);
    if (st) return st;
    CTLAB_REQUIRE(cap >= 0, "%s: cap < 0", fn);
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(counts && (cap == 0 || (idx && len)), "%s: output is null", fn);
    if constexpr (ND == 2) {
        if (mode == CTLAB_MODE_MARCH) {
            hipLaunchKernelGGL((k_trace<2, MODE_MARCH>), dim3(grid_for(n_ray)), dim3(kBlock), 0, stream, n_ray,
                               make_dims<2>(n), src, dst, minv, b, zero_rule, cap, idx, len, counts);
            return check_launch(fn);
        }
    }
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL((k_trace<ND, MODE_TORCH>), dim3(grid_for(n_ray)), dim3(kBlock), 0, stream, n_ray,
                           make_dims<ND>(n), src, dst, minv, b, zero_rule, cap, idx, len, counts);
    else
        hipLaunchKernelGGL((k_trace<ND, MODE_OTF>), dim3(grid_for(n_ray)), dim3(kBlock), 0, stream, n_ray,
                           make_dims<ND>(n), src, dst, minv, b, zero_rule, cap, idx, len, counts);
    return check_launch(fn);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int ctlab_forward_project_3d_otf(const float* volume, int64_t batch, int64_t n_x, int64_t n_y,
                                            int64_t n_z, const float* src, const float* dst, int64_t n_ray,
                                            const float* M_inv, const float* b, int mode, int64_t det_rows,
                                            int64_t det_cols, float* out, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return forward_otf<3>("forward_project_3d_otf", volume, batch, n, src, dst, n_ray, M_inv, b, mode,
                          CTLAB_ZERO_THRESH, det_rows, det_cols, out, as_stream(stream));
}

extern "C" int ctlab_back_project_3d_otf(const float* sino, int64_t batch, int64_t n_ray, const float* src,
                                         const float* dst, const float* M_inv, const float* b, int64_t n_x,
                                         int64_t n_y, int64_t n_z, int mode, int64_t det_rows, int64_t det_cols,
                                         int accumulate, float* out, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return back_otf<3>("back_project_3d_otf", sino, batch, n_ray, src, dst, M_inv, b, n, mode,
                       CTLAB_ZERO_THRESH, det_rows, det_cols, accumulate, out, as_stream(stream));
}

extern "C" int ctlab_forward_project_2d_otf(const float* image, int64_t batch, int64_t n_row, int64_t n_col,
                                            const float* src, const float* dst, int64_t n_ray,
                                            const float* M_inv, const float* b, int mode, int zero_rule,
                                            float* out, ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return forward_otf<2>("forward_project_2d_otf", image, batch, n, src, dst, n_ray, M_inv, b, mode, zero_rule,
                          0, 0, out, as_stream(stream));
}

extern "C" int ctlab_back_project_2d_otf(const float* sino, int64_t batch, int64_t n_ray, const float* src,
                                         const float* dst, const float* M_inv, const float* b, int64_t n_row,
                                         int64_t n_col, int mode, int zero_rule, int accumulate, float* out,
                                         ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return back_otf<2>("back_project_2d_otf", sino, batch, n_ray, src, dst, M_inv, b, n, mode, zero_rule, 0, 0,
                       accumulate, out, as_stream(stream));
}

extern "C" int ctlab_trace_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* src, const float* dst,
                              int64_t n_ray, const float* M_inv, const float* b, int mode, int zero_rule,
                              int64_t cap, int32_t* idx, float* len, int32_t* counts, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return trace<3>("trace_3d", n, src, dst, n_ray, M_inv, b, mode, zero_rule, cap, idx, len, counts,
                    as_stream(stream));
}

extern "C" int ctlab_trace_2d(int64_t n_row, int64_t n_col, const float* src, const float* dst, int64_t n_ray,
                              const float* M_inv, const float* b, int mode, int zero_rule, int64_t cap,
                              int32_t* idx, float* len, int32_t* counts, ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return trace<2>("trace_2d", n, src, dst, n_ray, M_inv, b, mode, zero_rule, cap, idx, len, counts,
                    as_stream(stream));
}
