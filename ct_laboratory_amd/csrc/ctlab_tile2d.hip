// ctlab_tile2d.hip -- atomic-free 2D back projection over image tiles, for any
// ray geometry (fan beam, static rings, parallel beams).
//
// A work unit is (tile, part): a 128x64 (or 64x64) tile of one image and a contiguous
// range of that scan's rays.  The unit's workgroup sums the range's
// contributions to the tile in LDS as 64-bit fixed point (exact and
// independent of the order of the rays, as in the 3D brick back projector)
// and then writes the tile once.  Per chunk of 4096 rays, a conservative slab
// test against the tile widened by the ray's rounding bound plus one pixel
// keeps the rays that can reach it, filed by estimated length into an LDS
// queue so that a wave's lanes trace rays of similar length; each kept ray is
// traced with the exact traversal clipped to the tile
// (trace_ray_window), so the tile receives exactly the segments whose pixel it
// contains.  Global float atomics, which execute past L2 and bound the
// one-ray-per-lane back projection, are left only for combining parts: a tile
// with one part is written with plain stores (deterministic); with several,
// each part adds each pixel once with a float atomic.
#include "ctlab_common.h"
#include "fixed_point.h"
#include "siddon_traverse.h"

namespace ctlab {

// Tiles are TR x 64 pixels, TR = 128 (66.5 KB of LDS) or 64 (33 KB), two
// workgroups per CU either way.  Larger tiles mean fewer (ray, tile) pairs to
// filter and set up: on the 64-scan batch 64 x 64 measured 10.0 ms, 128 x 64
// 7.2 ms, and 128 x 128 at one workgroup per CU 8.1.  The single-scan launch
// uses 128 x 64; the multi-scan launch takes the shape as an argument (tuned).
constexpr int kTileC = 64;                     // tile columns, pixels
constexpr int kTileStride = kTileC + 1;        // LDS row stride (64-bit cells), odd
constexpr int kTileThreads = 1024;             // 16 waves; two workgroups per CU
constexpr int kTilePerThread = 4;
constexpr int kTileChunk = kTilePerThread * kTileThreads;  // rays filtered per chunk (< 2^16)

static_assert(sizeof(ctlab_scan2d) == 88, "ctlab_scan2d layout");

// Rays of a chunk are traced in order of estimated crossing count, filed into
// quarter-octave buckets, so that a wave's lanes trace rays of similar length
// (in ray order a wave can span a tile's whole shadow, from corner-clipping to
// full-width rays).
constexpr int kTileBuckets = 32;

template <int TR>
struct TileLds {
    unsigned long long acc[TR * kTileStride];
    uint16_t queue[kTileChunk];
    float red[kTileThreads / 64];
    uint32_t count;
    uint32_t bcnt[kTileBuckets + 1];
};

// Longest segment any ray can have inside one pixel: the longer diagonal of
// the pixel parallelogram, |M d| for d in {(1,1), (1,-1)}, M = inverse of
// M_inv.  Evaluated in double.
__device__ __forceinline__ double pixel_diagonal(const Grid<2>& g) {
    const double a = g.minv[0], b = g.minv[1], c = g.minv[2], d = g.minv[3];
    const double det = a * d - b * c;
    const double p0 = (d - b) / det, p1 = (a - c) / det;  // M (1, 1)
    const double q0 = (d + b) / det, q1 = (-c - a) / det; // M (1, -1)
    return fmax(sqrt(p0 * p0 + p1 * p1), sqrt(q0 * q0 + q1 * q1));
}

// Conservative test: can ray s -> d have a segment whose rounded pixel lies in
// the window?  The traversal visits only segments whose rounded midpoint is in
// the window; a midpoint coordinate is within E of the exact one (E as in
// trace_ray_window, which drops the ray when E > 8), so such a ray passes
// through the window widened by E + 1/2.  The slab test here widens it by
// E + 1: the extra half pixel covers this test's own float rounding (relative
// 2^-22 of coordinates below 2^20 pixels).  It drops rays only for E > 8.5, so
// a rounding difference from the traversal's own E near 8 keeps the ray.
// Returns 0 for a ray that cannot reach the tile, else an estimate (>= 1) of
// its crossings in the widened tile.
template <int MODE, bool DIAG>
__device__ __forceinline__ uint32_t tile_may_hit(const Grid<2>& g, const Window<2>& win, const float* s,
                                                 const float* d) {
    float sv[2], dv[2];
    to_voxel_t<2, MODE, DIAG>(g, s, sv);
    to_voxel_t<2, MODE, DIAG>(g, d, dv);
    float t0 = 0.0f, t1 = 1.0f;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        float mag = fabsf(sv[a]) + fabsf(dv[a]) + (float)g.n[a] + 1.0f;
#pragma unroll
        for (int j = 0; j < 2; ++j) mag += fabsf(g.minv[a * 2 + j]) * (fabsf(s[j]) + fabsf(d[j]) + fabsf(g.b[j]));
        const float m = mag * 7.62939453125e-06f + 1.0f;  // E + 1 pixel, E = mag * 2^-17
        if (!(m <= 9.5f)) return 0u;  // E well above the traversal's limit of 8, or NaN: no segments
        const float lo = (float)win.lo[a] - 0.5f - m, hi = (float)(win.lo[a] + win.ext[a]) - 0.5f + m;
        const float dir = dv[a] - sv[a];
        if (fabsf(dir) < 1.0e-20f) {
            if (!(sv[a] >= lo && sv[a] <= hi)) return 0u;
            continue;
        }
        const float ta = (lo - sv[a]) / dir, tb = (hi - sv[a]) / dir;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    if (!(t0 <= t1)) return 0u;
    const float cross = (t1 - t0) * (fabsf(dv[0] - sv[0]) + fabsf(dv[1] - sv[1]));
    return (uint32_t)fminf(cross, 65535.0f) + 1u;
}

// Block-wide maximum (all kTileThreads threads call it).
__device__ __forceinline__ float tile_block_max(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    float r = red[0];
    for (int k = 1; k < kTileThreads / 64; ++k) r = fmaxf(r, red[k]);
    __syncthreads();  // red is reused
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r)));
}

// One work unit: rays [r0, r1) of a scan into tile `tile` of its image.
// direct: this part is the tile's only one (plain stores; else float atomics
// into an image that is zeroed or accumulated into).
template <int MODE, bool DIAG, int TR>
__device__ void tile_unit(const Grid<2>& g, const float* __restrict__ sn, const float* __restrict__ src,
                          const float* __restrict__ dst, float* __restrict__ img, int tile, int tiles_c, int64_t r0,
                          int64_t r1, bool direct, int accumulate, TileLds<TR>& L) {
    const int tid = threadIdx.x;
    Window<2> win;
    win.lo[0] = (tile / tiles_c) * TR;
    win.lo[1] = (tile % tiles_c) * kTileC;
    win.ext[0] = min(TR, g.n[0] - win.lo[0]);
    win.ext[1] = min(kTileC, g.n[1] - win.lo[1]);
    win.stride[0] = kTileStride;
    win.stride[1] = 1;
    // image index of tile pixel (r, c); pixels are enumerated as p = r * kTileC + c
    // over the full tile (shifts, not divisions), skipping those past a ragged edge
    auto gidx = [&](int r, int c) -> int64_t { return (int64_t)(win.lo[0] + r) * g.n[1] + (win.lo[1] + c); };

    for (int i = tid; i < TR * kTileStride; i += kTileThreads) L.acc[i] = 0ull;

    // 1. bound on |value| over the part's rays (INF when one is INF or NaN)
    float vmax = 0.0f;
    bool bad = false;
    for (int64_t r = r0 + tid; r < r1; r += kTileThreads) {
        const float a = fabsf(sn[r]);
        bad |= !(a <= 3.0e38f);
        vmax = fmaxf(vmax, a);
    }
    vmax = tile_block_max(bad ? __builtin_huge_valf() : vmax, L.red);

    // 2. power-of-two scale: a ray's segments in one pixel add up to at most
    // the pixel diagonal
    const FixedScale fs = fixed_scale(2.0 * (double)vmax * pixel_diagonal(g), (double)(r1 - r0), g.finite);

    if (!fs.fixed && direct && !accumulate) {  // float atomics below need a zeroed tile
        for (int p = tid; p < TR * kTileC; p += kTileThreads)
            if (p / kTileC < win.ext[0] && p % kTileC < win.ext[1]) img[gidx(p / kTileC, p % kTileC)] = 0.0f;
    }
    __syncthreads();

    // 3. filter, queue and trace, chunk by chunk
    if (vmax != 0.0f) {
        const float s1 = fs.s1, s2 = fs.s2;
        for (int64_t c0 = r0; c0 < r1; c0 += kTileChunk) {
            if (tid <= kTileBuckets) L.bcnt[tid] = 0u;
            __syncthreads();
            uint32_t key[kTilePerThread];
#pragma unroll
            for (int k = 0; k < kTilePerThread; ++k) {
                const int off = tid + k * kTileThreads;
                const int64_t r = c0 + off;
                key[k] = ~0u;
                if (r < r1 && sn[r] != 0.0f) {  // a zero value contributes exactly zero
                    const float s[2] = {src[2 * r], src[2 * r + 1]};
                    const float d[2] = {dst[2 * r], dst[2 * r + 1]};
                    const uint32_t est = tile_may_hit<MODE, DIAG>(g, win, s, d);
                    if (est) {
                        const uint32_t l2 = 31u - (uint32_t)__clz(est);  // est in [2^l2, 2^(l2+1))
                        const uint32_t bk = min(4u * l2 + ((est << 2) >> l2) - 4u, (uint32_t)(kTileBuckets - 1));
                        key[k] = (bk << 16) | atomicAdd(&L.bcnt[bk], 1u);
                    }
                }
            }
            __syncthreads();
            if (tid == 0) {
                uint32_t run = 0u;
                for (int k = 0; k < kTileBuckets; ++k) {
                    const uint32_t c = L.bcnt[k];
                    L.bcnt[k] = run;
                    run += c;
                }
                L.count = run;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kTilePerThread; ++k)
                if (key[k] != ~0u)
                    L.queue[L.bcnt[key[k] >> 16] + (key[k] & 0xffffu)] = (uint16_t)(tid + k * kTileThreads);
            __syncthreads();
            const uint32_t hits = L.count;
            for (uint32_t q = tid; q < hits; q += kTileThreads) {
                const int64_t r = c0 + L.queue[q];
                const float s[2] = {src[2 * r], src[2 * r + 1]};
                const float d[2] = {dst[2 * r], dst[2 * r + 1]};
                if (fs.fixed) {
                    const float val = sn[r] * s1 * s2;
                    trace_ray_window<2, MODE, DIAG, true>(
                        g, win, s, d, [&](uint32_t li, float len) { atomicAdd(&L.acc[li], to_fixed(val * len)); });
                } else {
                    const float val = sn[r];
                    trace_ray_window<2, MODE, DIAG, false>(g, win, s, d, [&](uint32_t li, float len) {
                        unsafeAtomicAdd(img + gidx((int)(li / kTileStride), (int)(li % kTileStride)), val * len);
                    });
                }
            }
            __syncthreads();  // queue and count are reused
        }
    }

    // 4. write the tile (rows of the image, coalesced)
    if (!fs.fixed) return;
    const double inv = ldexp(1.0, -fs.F);
    for (int p = tid; p < TR * kTileC; p += kTileThreads) {
        const int r = p / kTileC, c = p % kTileC;
        if (r >= win.ext[0] || c >= win.ext[1]) continue;
        const float v = (float)((double)(long long)L.acc[r * kTileStride + c] * inv);
        float* o = img + gidx(r, c);
        if (direct)
            *o = accumulate ? *o + v : v;
        else if (v != 0.0f)
            unsafeAtomicAdd(o, v);
    }
}

template <int MODE, int TR>
__device__ __forceinline__ void tile_unit_any(const Grid<2>& g, const float* sn, const float* src, const float* dst,
                                              float* img, int tile, int tiles_c, int64_t r0, int64_t r1, bool direct,
                                              int accumulate, TileLds<TR>& L) {
    if (g.diag)
        tile_unit<MODE, true, TR>(g, sn, src, dst, img, tile, tiles_c, r0, r1, direct, accumulate, L);
    else
        tile_unit<MODE, false, TR>(g, sn, src, dst, img, tile, tiles_c, r0, r1, direct, accumulate, L);
}

constexpr int kSingleScanTileRows = 128;

// One scan, batch elements along y: unit x = part * n_tiles + tile (the tiles
// of one ray range run together and share its rays in L2).
template <int MODE>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_back_tile2d(
    const float* __restrict__ sino, float* __restrict__ out, int64_t n_ray, const float* __restrict__ src,
    const float* __restrict__ dst, const float* __restrict__ minv, const float* __restrict__ b, int n_row, int n_col,
    int zero_rule, int tiles_c, int n_tiles, int64_t rays_per_part, int n_parts, int accumulate) {
    __shared__ TileLds<kSingleScanTileRows> L;
    const int n[2] = {n_row, n_col};
    const Grid<2> g = load_grid<2>(n, minv, b, zero_rule);
    const int part = blockIdx.x / n_tiles, tile = blockIdx.x % n_tiles;
    const int64_t r0 = (int64_t)part * rays_per_part, r1 = min(r0 + rays_per_part, n_ray);
    const int64_t bi = blockIdx.y;
    tile_unit_any<MODE, kSingleScanTileRows>(g, sino + bi * n_ray, src, dst, out + bi * (int64_t)n_row * n_col, tile,
                                             tiles_c, r0, r1,
                        n_parts == 1, accumulate, L);
}

// Many scans, persistent: workgroups take units from a counter.  Scan k owns
// units [item_begin_k, item_begin_k + tiles_k * parts_k), unit = part * tiles_k
// + tile, parts_k = ceil(n_ray_k / rays_per_part), tiles of TR x 64 pixels.
template <int TR>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch2d_tiles(
    const ctlab_scan2d* __restrict__ scans, int n_scan, int64_t total_units, int64_t rays_per_part, int accumulate,
    unsigned long long* __restrict__ counter) {
    __shared__ TileLds<TR> L;
    __shared__ int64_t s_unit;
    while (true) {
        if (threadIdx.x == 0) s_unit = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t unit = s_unit;
        __syncthreads();  // every lane has read s_unit before it is overwritten
        if (unit >= total_units) break;  // uniform: every workgroup reaches this exit
        int lo = 0, hi = n_scan;          // largest k with item_begin <= unit
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (scans[mid].item_begin <= unit) lo = mid; else hi = mid;
        }
        const ctlab_scan2d sc = scans[lo];
        const int n[2] = {sc.n_row, sc.n_col};
        const Grid<2> g = load_grid<2>(n, sc.minv, sc.b, sc.zero_rule);
        const int tiles_c = (sc.n_col + kTileC - 1) / kTileC;
        const int n_tiles = ((sc.n_row + TR - 1) / TR) * tiles_c;
        const int64_t local = unit - sc.item_begin;
        const int64_t part = local / n_tiles;
        const int tile = (int)(local % n_tiles);
        const int64_t r0 = part * rays_per_part, r1 = min(r0 + rays_per_part, sc.n_ray);
        if (r0 >= r1) continue;  // past the scan's units (a table with spare units): nothing to do
        const bool direct = sc.n_ray <= rays_per_part;
        const float* sn = reinterpret_cast<const float*>(sc.sino);
        const float* s = reinterpret_cast<const float*>(sc.src);
        const float* d = reinterpret_cast<const float*>(sc.dst);
        float* img = reinterpret_cast<float*>(sc.image);
        if (sc.mode == MODE_OTF)
            tile_unit_any<MODE_OTF, TR>(g, sn, s, d, img, tile, tiles_c, r0, r1, direct, accumulate, L);
        else
            tile_unit_any<MODE_TORCH, TR>(g, sn, s, d, img, tile, tiles_c, r0, r1, direct, accumulate, L);
        __syncthreads();  // the tile's LDS is cleared by the next unit
    }
}

// Single-scan 2D back projection (ctlab_back_project_2d_otf): sizes are
// checked by the caller.  Rays split into parts so that a launch has about
// 1024 units of at least one chunk each.
int back_project_2d_tiles(const char* fn, const float* sino, int64_t batch, int64_t n_ray, const float* src,
                          const float* dst, const float* minv, const float* b, const int64_t* n, int mode,
                          int zero_rule, int accumulate, float* out, hipStream_t stream) {
    const int tiles_c = (int)((n[1] + kTileC - 1) / kTileC);
    const int64_t n_tiles = ((n[0] + kSingleScanTileRows - 1) / kSingleScanTileRows) * tiles_c;
    CTLAB_REQUIRE(n_tiles <= (1LL << 24), "%s: image too large for the tile grid", fn);
    CTLAB_REQUIRE(n_ray >= 1, "%s: no rays", fn);
    int64_t parts = (1024 + n_tiles * batch - 1) / (n_tiles * batch);  // about 1024 units
    const int64_t max_parts = (n_ray + kTileChunk - 1) / kTileChunk;     // at least a chunk each
    if (parts > max_parts) parts = max_parts;
    if (parts < 1) parts = 1;
    const int64_t rpp = (n_ray + parts - 1) / parts;
    const int64_t n_parts = (n_ray + rpp - 1) / rpp;
    CTLAB_REQUIRE(n_parts * n_tiles <= kMaxGridX, "%s: too many work units", fn);
    if (n_parts > 1 && !accumulate) {
        int st = fill_f32(out, n[0] * n[1] * batch, 0.0f, stream);
        if (st) return st;
    }
    dim3 grid((unsigned)(n_parts * n_tiles), (unsigned)batch);
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL(k_back_tile2d<MODE_TORCH>, grid, dim3(kTileThreads), 0, stream, sino, out, n_ray, src, dst,
                           minv, b, (int)n[0], (int)n[1], zero_rule, tiles_c, (int)n_tiles, rpp, (int)n_parts,
                           accumulate);
    else
        hipLaunchKernelGGL(k_back_tile2d<MODE_OTF>, grid, dim3(kTileThreads), 0, stream, sino, out, n_ray, src, dst,
                           minv, b, (int)n[0], (int)n[1], zero_rule, tiles_c, (int)n_tiles, rpp, (int)n_parts,
                           accumulate);
    return check_launch(fn);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int64_t ctlab_back_project_2d_tile_units(int64_t n_row, int64_t n_col, int64_t n_ray,
                                                     int64_t rays_per_part, int64_t tile_rows) {
    if (n_row < 1 || n_col < 1 || n_ray < 0 || rays_per_part < 1 || (tile_rows != 64 && tile_rows != 128)) return -1;
    const int64_t tiles = ((n_row + tile_rows - 1) / tile_rows) * ((n_col + kTileC - 1) / kTileC);
    return tiles * ((n_ray + rays_per_part - 1) / rays_per_part);
}

extern "C" int ctlab_back_project_2d_batch_tiled(const ctlab_scan2d* scans, int64_t n_scan, int64_t total_units,
                                                 int64_t rays_per_part, int64_t tile_rows, int64_t n_groups,
                                                 void* counter, int accumulate, ctlab_stream_t stream) {
    const char* fn = "back_project_2d_batch_tiled";
    CTLAB_REQUIRE(n_scan >= 0 && n_scan <= 0x7fffffffLL, "%s: n_scan out of range", fn);
    CTLAB_REQUIRE(total_units >= 0 && rays_per_part >= 1 && rays_per_part <= (1LL << 40), "%s: bad work units", fn);
    CTLAB_REQUIRE(tile_rows == 64 || tile_rows == 128, "%s: tile_rows %lld (64 or 128)", fn, (long long)tile_rows);
    CTLAB_REQUIRE(n_groups >= 1 && n_groups <= 65536, "%s: n_groups %lld out of range", fn, (long long)n_groups);
    if (n_scan == 0) return CTLAB_OK;
    CTLAB_REQUIRE(scans && counter, "%s: null pointer", fn);
    auto* ctr = reinterpret_cast<unsigned long long*>(counter);
    int st = batch2d_clear(fn, scans, n_scan, !accumulate, ctr, as_stream(stream));
    if (st || total_units == 0) return st;
    if (tile_rows == 128)
        hipLaunchKernelGGL(k_batch2d_tiles<128>, dim3((unsigned)n_groups), dim3(kTileThreads), 0, as_stream(stream),
                           scans, (int)n_scan, total_units, rays_per_part, accumulate, ctr);
    else
        hipLaunchKernelGGL(k_batch2d_tiles<64>, dim3((unsigned)n_groups), dim3(kTileThreads), 0, as_stream(stream),
                           scans, (int)n_scan, total_units, rays_per_part, accumulate, ctr);
    return check_launch(fn);
}
