// ctlab_table_brick.hip -- back projection over precomputed crossing tables
// (float, or uint16 with the CUDA decode) that accumulates in LDS bricks
// instead of adding every segment to the volume with a global atomic
// (the reference's consumers, src/ct_projector_3d.cu:215-269 and 325-376, add
// each segment with atomicAdd).
//
// A table fixes each ray's segments and their voxels.  The volume is cut into
// bricks of 32 x 32 x 16 voxels (64 KB of floats).  A geometry-only pass
// splits every ray's row into runs: maximal stretches of consecutive visited
// segments whose voxels lie in one brick (a ray that comes back to a brick
// starts another run).  Runs are grouped by brick; a work unit is a slice of
// one brick's runs.  The back projection gives each unit one block, which
// walks its runs' segments exactly as the table walkers do (same entries, same
// arithmetic, same skip rules), adds val * len into the LDS brick with LDS
// atomics, and then adds the brick's nonzero voxels to the volume with one
// global atomic each.  Every segment the reference's consumer adds is added
// once, to the same voxel, with the same value; only the grouping of the
// float sums differs (the reference's atomic order is unspecified as well).
//
//   count   runs per brick (global atomics on the brick counters)
//   fill    run records placed at offsets[brick] + a per-brick cursor
//   back    one block per (unit, batch item)
#include "ctlab_common.h"
#include "siddon_traverse.h"
#include "table_walk.h"

#include <type_traits>

namespace ctlab {

constexpr int kTbX = CTLAB_TABLE_BRICK_X, kTbY = CTLAB_TABLE_BRICK_Y, kTbZ = CTLAB_TABLE_BRICK_Z;
constexpr int kTbVox = kTbX * kTbY * kTbZ;
constexpr int kTbBlock = 512;  // 8 waves per block, two blocks (128 KB of LDS) per CU

static_assert(sizeof(ctlab_table_run_t) == 16, "run record is 16 bytes");
static_assert(sizeof(ctlab_table_unit_t) == 16, "unit record is 16 bytes");

struct Dims3 {
    int n[3];
};

struct BrickGrid {
    int nb[3];  // bricks per axis
    __host__ __device__ __forceinline__ uint32_t id(const int32_t* r) const {
        return ((uint32_t)(r[0] / kTbX) * (uint32_t)nb[1] + (uint32_t)(r[1] / kTbY)) * (uint32_t)nb[2] +
               (uint32_t)(r[2] / kTbZ);
    }
};

// Pass over every ray's row: FILL = false counts the runs of each brick,
// FILL = true writes each run's record at offsets[brick] + cursor[brick]++.
template <bool U16, bool FILL>
__global__ __launch_bounds__(kBlock) void k_table_runs(Dims3 dims, BrickGrid bg, const float* __restrict__ tvals,
                                                       const uint16_t* __restrict__ u16,
                                                       const float* __restrict__ start,
                                                       const float* __restrict__ scale, int64_t K,
                                                       const float* __restrict__ src, const float* __restrict__ dst,
                                                       int64_t n_ray, const float* __restrict__ minv,
                                                       const float* __restrict__ b, uint32_t* __restrict__ counts,
                                                       const int64_t* __restrict__ offsets,
                                                       ctlab_table_run_t* __restrict__ runs, int64_t n_runs) {
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t r = r0 + threadIdx.x;
    const bool active = r < n_ray;
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    float s[3] = {}, d[3] = {};
    if (active) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s[a] = src[r * 3 + a];
            d[a] = dst[r * 3 + a];
        }
    }
    uint32_t cur = 0xffffffffu, k0 = 0;
    int64_t seg0 = 0, last = 0;
    auto close = [&]() {
        if (!FILL) {
            atomicAdd(counts + cur, 1u);
        } else {
            const int64_t pos = offsets[cur] + (int64_t)atomicAdd(counts + cur, 1u);
            if (pos < n_runs) {
                ctlab_table_run_t rec;
                rec.ray = (uint32_t)r;
                rec.seg0 = (uint32_t)seg0;
                rec.nseg = (uint32_t)(last - seg0 + 1);
                rec.k0 = k0;
                runs[pos] = rec;
            }
        }
    };
    auto visit = [&](int64_t j, uint32_t kp, const float* c, float) {
        int32_t v[3];
        if (!round_voxel<3>(g, c, v)) return;
        const uint32_t id = bg.id(v);
        if (id != cur) {
            if (cur != 0xffffffffu) close();
            cur = id;
            seg0 = j;
            k0 = kp;
        }
        last = j;
    };
    if constexpr (U16) {
        __shared__ RowStage<uint16_t, kStageU16> st;
        const float st0 = active ? start[r] : 0.0f, sc = active ? scale[r] : 0.0f;
        if (g.diag)
            walk_u16<true>(g, s, d, u16, K, r0, n_ray, active, st0, sc, st, visit);
        else
            walk_u16<false>(g, s, d, u16, K, r0, n_ray, active, st0, sc, st, visit);
    } else {
        __shared__ RowStage<float, kStageF32> st;
        if (g.diag)
            walk_table<3, true>(g, s, d, tvals, K, r0, n_ray, active, st, visit);
        else
            walk_table<3, false>(g, s, d, tvals, K, r0, n_ray, active, st, visit);
    }
    if (active && cur != 0xffffffffu) close();
}

// A table block in flight: raw is the 16-byte load, fix the block's entries
// when it is not wholly inside the table (zero outside it), taken from the
// table's two edge blocks staged in LDS.  raw is then loaded from a block that
// is wholly inside (first_full, the first entry of one), so every block costs
// one global load and the value is selected only when the block is used: a
// block loaded ahead is not waited for at the load, as it would be if a path
// with loads of its own wrote into the load's registers.  A table with no
// whole block (first_full < 0) takes its entries one by one.
template <typename V>
struct TableBlock {
    V raw, fix;
    bool inside;
    __device__ __forceinline__ V value() const { return inside ? raw : fix; }
};

// Block q of 8 uint16 entries: entries [8q - off, 8q - off + 8) of the flat
// table of n_ent entries, whose first entry sits `off` entries past a 16-byte
// boundary.  edge holds blocks 0 and (n_ent - 1 + off) / 8; blocks past the
// table are never consumed.
__device__ __forceinline__ TableBlock<uint4> load_block(const uint16_t* __restrict__ u16, int64_t q, int off,
                                                        int64_t n_ent, int64_t first_full, const uint4* edge) {
    TableBlock<uint4> b;
    const int64_t e0 = q * 8 - off;
    b.inside = e0 >= 0 && e0 + 8 <= n_ent;
    if (first_full < 0) {  // uniform: a table with no whole block
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        for (int i = 0; i < 8; ++i) {
            const int64_t e = e0 + i;
            if (e >= 0 && e < n_ent) w[i >> 1] |= (uint32_t)u16[e] << ((i & 1) * 16);
        }
        b.raw = b.fix = make_uint4(w[0], w[1], w[2], w[3]);
        return b;
    }
    b.raw = *reinterpret_cast<const uint4*>(u16 + (b.inside ? e0 : first_full));
    b.fix = (e0 < 0) ? edge[0] : edge[1];
    return b;
}

// Block q of 4 float entries: entries [4q - off, 4q - off + 4) of the flat
// table of n_ent entries, whose first entry sits `off` entries past a 16-byte
// boundary; loaded as load_block loads uint16 blocks.
__device__ __forceinline__ TableBlock<float4> load_block4(const float* __restrict__ tv, int64_t q, int off,
                                                          int64_t n_ent, int64_t first_full, const float4* edge) {
    TableBlock<float4> b;
    const int64_t e0 = q * 4 - off;
    b.inside = e0 >= 0 && e0 + 4 <= n_ent;
    if (first_full < 0) {
        float w[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 4; ++i) {
            const int64_t e = e0 + i;
            if (e >= 0 && e < n_ent) w[i] = tv[e];
        }
        b.raw = b.fix = make_float4(w[0], w[1], w[2], w[3]);
        return b;
    }
    b.raw = *reinterpret_cast<const float4*>(tv + (b.inside ? e0 : first_full));
    b.fix = (e0 < 0) ? edge[0] : edge[1];
    return b;
}

// First entry of the first block wholly inside a table of n_ent entries whose
// first entry sits `off` entries (of per_block per block) past a block
// boundary, or -1 when there is none.
__device__ __forceinline__ int64_t first_full_block(int64_t n_ent, int off, int per_block) {
    const int64_t e = (off == 0) ? 0 : per_block - off;
    return (e + per_block <= n_ent) ? e : -1;
}

__device__ __forceinline__ float pick4(const float4& b, int64_t gi) {
    return (gi & 2) ? ((gi & 1) ? b.w : b.z) : ((gi & 1) ? b.y : b.x);
}

// The runs of one unit added into the LDS brick acc (origin x0, y0, z0).
template <bool U16, bool DIAG>
__device__ __forceinline__ void table_brick_runs(float* acc, const Grid<3>& g, int x0, int y0, int z0,
                                                 const float* __restrict__ sino, const float* __restrict__ tvals,
                                                 const uint16_t* __restrict__ u16, const float* __restrict__ start,
                                                 const float* __restrict__ scale, int64_t K,
                                                 const float* __restrict__ src, const float* __restrict__ dst,
                                                 int64_t n_ray, const ctlab_table_run_t* __restrict__ runs,
                                                 int64_t n_runs, const ctlab_table_unit_t& u, int64_t bi,
                                                 int u16_off, int f32_off, const uint4* edge16,
                                                 const float4* edge32) {
    MidIndex<3, DIAG> mi;
    mi.init(g);
    const float* __restrict__ sn = sino + bi * n_ray;
    auto add = [&](const float* c, float w) {
        int32_t v[3];
        if (!round_voxel<3>(g, c, v)) return;
        const uint32_t lx = (uint32_t)(v[0] - x0), ly = (uint32_t)(v[1] - y0), lz = (uint32_t)(v[2] - z0);
        if (lx < (uint32_t)kTbX && ly < (uint32_t)kTbY && lz < (uint32_t)kTbZ)
            atomicAdd(&acc[(lx * kTbY + ly) * kTbZ + lz], w);
    };
    // A lane's runs are q = tid, tid + kTbBlock, ...: the next run's record is
    // loaded while the current run is walked, and a run's value, ray and table
    // loads are all issued before its value is tested.
    const int64_t n_q = min((int64_t)u.n, n_runs - u.first);
    int64_t q = threadIdx.x;
    ctlab_table_run_t nrec = {};
    if (q < n_q) nrec = runs[u.first + q];
    for (; q < n_q; q += kTbBlock) {
        const ctlab_table_run_t rr = nrec;
        if (q + kTbBlock < n_q) nrec = runs[u.first + q + kTbBlock];
        const int64_t ray = rr.ray;
        const int64_t j0 = rr.seg0, j1 = (int64_t)rr.seg0 + rr.nseg;
        if (ray >= n_ray || j0 < 1 || j1 > K) continue;
        const float val = sn[ray];
        float s[3], vw[3], p0[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s[a] = src[ray * 3 + a];
            vw[a] = dst[ray * 3 + a] - s[a];
        }
        if (U16) {
            const float st0 = start[ray], sc = scale[ray];
            // entries read 8 at a time in 16-byte blocks (gi counts entries from the
            // 16-byte boundary at or before the table's start); the next block is
            // loaded while the current one is walked
            const int64_t n_ent = n_ray * K;
            int64_t gi = ray * K + j0 + u16_off;
            const int64_t ff = first_full_block(n_ent, u16_off, 8);
            const TableBlock<uint4> first = load_block(u16, gi >> 3, u16_off, n_ent, ff, edge16);
            TableBlock<uint4> nxt = load_block(u16, (gi >> 3) + 1, u16_off, n_ent, ff, edge16);
            if (val == 0.0f) continue;
            // t of entry j0-1: start for entry 0, else start + k * scale (3d.cu:128-129)
            const float t0 = (j0 == 1) ? st0 : __builtin_fmaf((float)rr.k0, sc, st0);
#pragma unroll
            for (int a = 0; a < 3; ++a) p0[a] = s[a] + t0 * vw[a];
            uint32_t ks = rr.k0;
            uint4 blk = first.value();
            for (int64_t j = j0; j < j1; ++j, ++gi) {
                if ((gi & 7) == 0 && j != j0) {
                    blk = nxt.value();
                    nxt = load_block(u16, (gi >> 3) + 1, u16_off, n_ent, ff, edge16);
                }
                const uint32_t w = (gi & 4) ? ((gi & 2) ? blk.w : blk.z) : ((gi & 2) ? blk.y : blk.x);
                ks += (gi & 1) ? (w >> 16) : (w & 0xffffu);
                const float t1 = __builtin_fmaf((float)ks, sc, st0);
                float c[3];
                const float len = segment_mid<3, DIAG>(mi, s, vw, p0, t1, c);
                if (!(len >= 1e-9f)) continue;  // 3d.cu:137
                add(c, val * len);
            }
        } else {
            // entries read 4 at a time in 16-byte blocks, as the uint16 ones
            const int64_t n_ent = n_ray * K;
            int64_t gi = ray * K + j0 - 1 + f32_off;
            const int64_t ff = first_full_block(n_ent, f32_off, 4);
            float4 blk = load_block4(tvals, gi >> 2, f32_off, n_ent, ff, edge32).value();
            TableBlock<float4> nxt = load_block4(tvals, (gi >> 2) + 1, f32_off, n_ent, ff, edge32);
            if (val == 0.0f) continue;
            float t0 = pick4(blk, gi);
#pragma unroll
            for (int a = 0; a < 3; ++a) p0[a] = s[a] + t0 * vw[a];
            for (int64_t j = j0; j < j1; ++j) {
                ++gi;
                if ((gi & 3) == 0) {
                    blk = nxt.value();
                    nxt = load_block4(tvals, (gi >> 2) + 1, f32_off, n_ent, ff, edge32);
                }
                const float t1 = pick4(blk, gi);
                if (isinf(t0) || isinf(t1)) {  // end of a sorted row; a finite value restarts the walk
                    t0 = t1;
#pragma unroll
                    for (int a = 0; a < 3; ++a) p0[a] = s[a] + t0 * vw[a];
                    continue;
                }
                float c[3];
                const float len = segment_mid<3, DIAG>(mi, s, vw, p0, t1, c);
                t0 = t1;
                if (len > 0.0f) add(c, val * len);
            }
        }
    }
}

template <bool U16>
__global__ __launch_bounds__(kTbBlock) void k_back_table_brick(
    const float* __restrict__ sino, Dims3 dims, BrickGrid bg, const float* __restrict__ tvals,
    const uint16_t* __restrict__ u16, const float* __restrict__ start, const float* __restrict__ scale, int64_t K,
    const float* __restrict__ src, const float* __restrict__ dst, int64_t n_ray, const float* __restrict__ minv,
    const float* __restrict__ b, const ctlab_table_run_t* __restrict__ runs, int64_t n_runs,
    const ctlab_table_unit_t* __restrict__ units, int64_t vol_elems, float* __restrict__ out, int u16_off,
    int f32_off) {
    __shared__ float acc[kTbVox];
    const ctlab_table_unit_t u = units[blockIdx.x];
    const int64_t bi = blockIdx.y;
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const uint32_t n_brick = (uint32_t)bg.nb[0] * (uint32_t)bg.nb[1] * (uint32_t)bg.nb[2];
    if (u.brick >= n_brick || !g.finite) return;  // uniform over the block
    const int bz = (int)(u.brick % (uint32_t)bg.nb[2]);
    const int by = (int)((u.brick / (uint32_t)bg.nb[2]) % (uint32_t)bg.nb[1]);
    const int bx = (int)(u.brick / ((uint32_t)bg.nb[2] * (uint32_t)bg.nb[1]));
    const int x0 = bx * kTbX, y0 = by * kTbY, z0 = bz * kTbZ;
    for (int i = threadIdx.x; i < kTbVox; i += kTbBlock) acc[i] = 0.0f;
    // the table's first and last 16-byte blocks (entries outside the table: 0)
    __shared__ __align__(16) uint32_t edge[8];
    {
        const int64_t n_ent = n_ray * K;
        const int t = threadIdx.x;
        if (U16 && t < 16) {
            const int off = u16_off;
            const int64_t q = (t < 8) ? 0 : (n_ent - 1 + off) / 8;
            const int64_t e = q * 8 - off + (t & 7);
            reinterpret_cast<uint16_t*>(edge)[t] = (e >= 0 && e < n_ent) ? u16[e] : (uint16_t)0;
        }
        if (!U16 && t < 8) {
            const int off = f32_off;
            const int64_t q = (t < 4) ? 0 : (n_ent - 1 + off) / 4;
            const int64_t e = q * 4 - off + (t & 3);
            reinterpret_cast<float*>(edge)[t] = (e >= 0 && e < n_ent) ? tvals[e] : 0.0f;
        }
    }
    __syncthreads();
    if (g.diag)
        table_brick_runs<U16, true>(acc, g, x0, y0, z0, sino, tvals, u16, start, scale, K, src, dst, n_ray, runs,
                                    n_runs, u, bi, u16_off, f32_off, reinterpret_cast<const uint4*>(edge),
                                    reinterpret_cast<const float4*>(edge));
    else
        table_brick_runs<U16, false>(acc, g, x0, y0, z0, sino, tvals, u16, start, scale, K, src, dst, n_ray, runs,
                                     n_runs, u, bi, u16_off, f32_off, reinterpret_cast<const uint4*>(edge),
                                    reinterpret_cast<const float4*>(edge));
    __syncthreads();
    float* __restrict__ o = out + bi * vol_elems;
    for (int i = threadIdx.x; i < kTbVox; i += kTbBlock) {
        const float w = acc[i];
        if (w == 0.0f) continue;
        const int x = x0 + i / (kTbY * kTbZ), y = y0 + (i / kTbZ) % kTbY, z = z0 + i % kTbZ;
        if (x < dims.n[0] && y < dims.n[1] && z < dims.n[2])
            unsafeAtomicAdd(o + ((int64_t)x * dims.n[1] + y) * dims.n[2] + z, w);
    }
}

// Forward projection from a table over each ray's visited range only: segments
// first[r] .. end[r]-1, with k_first[r] the uint16 running sum at entry
// first[r]-1 (0, 0 = no visited segment).  Outside that range no segment is
// visited, so the sum -- the same fused multiply-adds over the same segments
// in the same order as k_forward_u16 / k_forward_tvals -- is bit for bit
// theirs.  The block stages only the row chunks its rays' ranges cover.
template <bool U16>
__global__ __launch_bounds__(kBlock) void k_forward_table_ranged(
    const float* __restrict__ vol, float* __restrict__ out, int64_t n_ray, int64_t vol_elems, Dims3 dims,
    const float* __restrict__ tvals, const uint16_t* __restrict__ u16, const float* __restrict__ start,
    const float* __restrict__ scale, int64_t K, const int32_t* __restrict__ first, const int32_t* __restrict__ end,
    const uint32_t* __restrict__ k_first, const float* __restrict__ src, const float* __restrict__ dst,
    const float* __restrict__ minv, const float* __restrict__ b) {
    using T = typename std::conditional<U16, uint16_t, float>::type;
    constexpr int C = U16 ? kStageU16 : kStageF32;
    __shared__ RowStage<T, C> st;
    __shared__ int s_lo, s_hi;
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t r = r0 + threadIdx.x;
    const bool active = r < n_ray;
    const int64_t bi = blockIdx.y;
    const int64_t jf = active ? first[r] : 0, je = active ? end[r] : 0;
    const bool walk = active && jf >= 1 && jf < je && je <= K;
    if (threadIdx.x == 0) {
        s_lo = 0x7fffffff;
        s_hi = 0;
    }
    __syncthreads();
    if (walk) {
        atomicMin(&s_lo, (int)jf);
        atomicMax(&s_hi, (int)je);
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;  // block-uniform
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    const float* __restrict__ v = vol + bi * vol_elems;
    // gathers consumed one segment later (their loads overlap the next
    // segment); the same chain of fused multiply-adds, fma(0, 0, +0) = +0 first
    float acc = 0.0f, pv = 0.0f, pl = 0.0f;
    if (lo < hi && g.finite) {
        float s[3] = {}, vw[3] = {}, p0[3] = {};
        float st0 = 0.0f, sc = 0.0f, t0 = 0.0f;
        uint32_t ks = 0;
        if (walk) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s[a] = src[r * 3 + a];
                vw[a] = dst[r * 3 + a] - s[a];
            }
            if (U16) {
                st0 = start[r];
                sc = scale[r];
                ks = k_first[r];
                // t of entry first-1: start for entry 0, else start + k * scale (3d.cu:128-129)
                t0 = (jf == 1) ? st0 : __builtin_fmaf((float)ks, sc, st0);
            } else {
                t0 = tvals[r * K + jf - 1];
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) p0[a] = s[a] + t0 * vw[a];
        }
        auto run = [&](auto diag) {
            constexpr bool DIAG = decltype(diag)::value;
            MidIndex<3, DIAG> mi;
            mi.init(g);
            const T* __restrict__ mine = st.e + threadIdx.x * RowStage<T, C>::kStride;
            const T* __restrict__ table;
            if constexpr (U16) table = u16; else table = tvals;
            for (int64_t c0 = (int64_t)(lo / C) * C; c0 < hi; c0 += C) {
                stage_rows(st, table, K, r0, n_ray, c0);
                if (!walk) continue;
                const int64_t j0 = max(c0, jf), j1 = min(min(c0 + C, je), K);
                for (int64_t j = j0; j < j1; ++j) {
                    float c[3];
                    if constexpr (U16) {
                        const uint32_t dk = (uint32_t)mine[j - c0];
                        if (dk == 0u) continue;  // same t: a segment of length 0
                        ks += dk;
                        const float t1 = __builtin_fmaf((float)ks, sc, st0);
                        const float len = segment_mid<3, DIAG>(mi, s, vw, p0, t1, c);
                        if (!(len >= 1e-9f)) continue;  // 3d.cu:137
                        uint32_t idx;
                        if (flat_voxel<3>(g, c, &idx)) {
                            acc = __builtin_fmaf(pv, pl, acc);
                            pv = v[idx];
                            pl = len;
                        }
                    } else {
                        const float t1 = mine[j - c0];
                        if (isinf(t0) || isinf(t1)) {  // end of a sorted row; a finite value restarts the walk
                            t0 = t1;
#pragma unroll
                            for (int a = 0; a < 3; ++a) p0[a] = s[a] + t0 * vw[a];
                            continue;
                        }
                        const float len = segment_mid<3, DIAG>(mi, s, vw, p0, t1, c);
                        t0 = t1;
                        uint32_t idx;
                        if (len > 0.0f && flat_voxel<3>(g, c, &idx)) {
                            acc = __builtin_fmaf(pv, pl, acc);
                            pv = v[idx];
                            pl = len;
                        }
                    }
                }
            }
        };
        if (g.diag)
            run(std::integral_constant<bool, true>());
        else
            run(std::integral_constant<bool, false>());
    }
    acc = __builtin_fmaf(pv, pl, acc);
    if (active) out[bi * n_ray + r] = acc;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static BrickGrid brick_grid(const int64_t* n) {
    BrickGrid bg;
    bg.nb[0] = (int)((n[0] + kTbX - 1) / kTbX);
    bg.nb[1] = (int)((n[1] + kTbY - 1) / kTbY);
    bg.nb[2] = (int)((n[2] + kTbZ - 1) / kTbZ);
    return bg;
}

struct TableArgs {
    const float* tvals;
    const uint16_t* u16;
    const float* start;
    const float* scale;
    int64_t K;
};

static int check_table_geometry(const char* fn, const int64_t* n, const TableArgs& t, const float* src,
                                const float* dst, int64_t n_ray, const float* minv, const float* b) {
    for (int a = 0; a < 3; ++a)
        CTLAB_REQUIRE(n[a] >= 1 && n[a] <= 0x7fffffffLL, "%s: grid dimension %d is %lld (must be >= 1)", fn, a,
                      (long long)n[a]);
    CTLAB_REQUIRE(n[0] * n[1] * n[2] <= kMaxVoxels, "%s: grid has more than 2^31-1 voxels", fn);
    int st = check_index_limits(fn, n, 3);
    if (st) return st;
    CTLAB_REQUIRE((t.tvals != nullptr) != (t.u16 != nullptr), "%s: pass exactly one of tvals and u16", fn);
    CTLAB_REQUIRE(!t.u16 || (t.start && t.scale), "%s: a uint16 table needs start and scale", fn);
    CTLAB_REQUIRE(t.K >= 0 && t.K <= 0x7fffffffLL, "%s: K %lld out of range", fn, (long long)t.K);
    CTLAB_REQUIRE(n_ray >= 0 && n_ray <= 0xffffffffLL, "%s: n_ray %lld out of range [0, 2^32)", fn,
                  (long long)n_ray);
    CTLAB_REQUIRE(grid_for(n_ray) <= kMaxGridX, "%s: too many rays", fn);
    if (n_ray > 0) CTLAB_REQUIRE(src && dst && minv && b, "%s: null pointer", fn);
    return CTLAB_OK;
}

template <bool FILL>
static int table_runs(const char* fn, const int64_t* n, const TableArgs& t, const float* src, const float* dst,
                      int64_t n_ray, const float* minv, const float* b, uint32_t* counts, const int64_t* offsets,
                      ctlab_table_run_t* runs, int64_t n_runs, hipStream_t stream) {
    int st = check_table_geometry(fn, n, t, src, dst, n_ray, minv, b);
    if (st) return st;
    const BrickGrid bg = brick_grid(n);
    const int64_t n_brick = (int64_t)bg.nb[0] * bg.nb[1] * bg.nb[2];
    CTLAB_REQUIRE(counts, "%s: counts is null", fn);
    if (FILL) CTLAB_REQUIRE(offsets && (n_runs == 0 || runs) && n_runs >= 0, "%s: null run buffers", fn);
    // zero bits: the counters are cleared by the fill kernel (never hipMemsetAsync)
    st = fill_f32(reinterpret_cast<float*>(counts), n_brick, 0.0f, stream);
    if (st) return st;
    if (n_ray == 0 || t.K < 2) return CTLAB_OK;
    Dims3 dims;
    for (int a = 0; a < 3; ++a) dims.n[a] = (int)n[a];
    const dim3 grid(grid_for(n_ray)), block(kBlock);
    if (t.u16)
        hipLaunchKernelGGL((k_table_runs<true, FILL>), grid, block, 0, stream, dims, bg, t.tvals, t.u16, t.start,
                           t.scale, t.K, src, dst, n_ray, minv, b, counts, offsets, runs, n_runs);
    else
        hipLaunchKernelGGL((k_table_runs<false, FILL>), grid, block, 0, stream, dims, bg, t.tvals, t.u16, t.start,
                           t.scale, t.K, src, dst, n_ray, minv, b, counts, offsets, runs, n_runs);
    return check_launch(fn);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int64_t ctlab_table_brick_count(int64_t n_x, int64_t n_y, int64_t n_z) {
    if (n_x < 1 || n_y < 1 || n_z < 1) return 0;
    const int64_t n[3] = {n_x, n_y, n_z};
    const BrickGrid bg = brick_grid(n);
    return (int64_t)bg.nb[0] * bg.nb[1] * bg.nb[2];
}

extern "C" int ctlab_table_runs_count_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* tvals,
                                         const uint16_t* u16, const float* start, const float* scale, int64_t K,
                                         const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                         const float* b, uint32_t* counts, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return table_runs<false>("ctlab_table_runs_count_3d", n, TableArgs{tvals, u16, start, scale, K}, src, dst, n_ray,
                             M_inv, b, counts, nullptr, nullptr, 0, as_stream(stream));
}

extern "C" int ctlab_table_runs_fill_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* tvals,
                                        const uint16_t* u16, const float* start, const float* scale, int64_t K,
                                        const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                        const float* b, const int64_t* offsets, uint32_t* cursor,
                                        ctlab_table_run_t* runs, int64_t n_runs, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return table_runs<true>("ctlab_table_runs_fill_3d", n, TableArgs{tvals, u16, start, scale, K}, src, dst, n_ray,
                            M_inv, b, cursor, offsets, runs, n_runs, as_stream(stream));
}

extern "C" int ctlab_back_project_3d_table_brick(const float* sino, int64_t batch, const float* tvals,
                                                 const uint16_t* u16, const float* start, const float* scale,
                                                 int64_t K, const float* src, const float* dst, int64_t n_ray,
                                                 const float* M_inv, const float* b, int64_t n_x, int64_t n_y,
                                                 int64_t n_z, const ctlab_table_run_t* runs, int64_t n_runs,
                                                 const ctlab_table_unit_t* units, int64_t n_units, int accumulate,
                                                 float* out, ctlab_stream_t stream) {
    static const char* fn = "ctlab_back_project_3d_table_brick";
    const int64_t n[3] = {n_x, n_y, n_z};
    const TableArgs t{tvals, u16, start, scale, K};
    int st = check_table_geometry(fn, n, t, src, dst, n_ray, M_inv, b);
    if (st) return st;
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range [1, 65535]", fn, (long long)batch);
    CTLAB_REQUIRE(n_runs >= 0 && n_units >= 0 && n_units <= kMaxGridX, "%s: run/unit counts out of range", fn);
    CTLAB_REQUIRE(out, "%s: out is null", fn);
    const int64_t V = n_x * n_y * n_z;
    hipStream_t s = as_stream(stream);
    if (!accumulate) {
        st = fill_f32(out, V * batch, 0.0f, s);
        if (st) return st;
    }
    if (n_ray == 0 || n_units == 0 || K < 2) return CTLAB_OK;
    CTLAB_REQUIRE(sino && runs && units, "%s: null pointer", fn);
    CTLAB_REQUIRE(!u16 || ((uintptr_t)u16 & 1) == 0, "%s: the uint16 table is not 2-byte aligned", fn);
    CTLAB_REQUIRE(!tvals || ((uintptr_t)tvals & 3) == 0, "%s: the float table is not 4-byte aligned", fn);
    const int u16_off = u16 ? (int)(((uintptr_t)u16 & 15) >> 1) : 0;
    const int f32_off = tvals ? (int)(((uintptr_t)tvals & 15) >> 2) : 0;
    Dims3 dims;
    for (int a = 0; a < 3; ++a) dims.n[a] = (int)n[a];
    const BrickGrid bg = brick_grid(n);
    const dim3 grid((unsigned)n_units, (unsigned)batch), block(kTbBlock);
    if (u16)
        hipLaunchKernelGGL((k_back_table_brick<true>), grid, block, 0, s, sino, dims, bg, tvals, u16, start, scale, K,
                           src, dst, n_ray, M_inv, b, runs, n_runs, units, V, out, u16_off, 0);
    else
        hipLaunchKernelGGL((k_back_table_brick<false>), grid, block, 0, s, sino, dims, bg, tvals, u16, start, scale, K,
                           src, dst, n_ray, M_inv, b, runs, n_runs, units, V, out, 0, f32_off);
    return check_launch(fn);
}

extern "C" int ctlab_forward_project_3d_table_ranged(const float* volume, int64_t batch, int64_t n_x, int64_t n_y,
                                                    int64_t n_z, const float* tvals, const uint16_t* u16,
                                                    const float* start, const float* scale, int64_t K,
                                                    const int32_t* first, const int32_t* end,
                                                    const uint32_t* k_first, const float* src, const float* dst,
                                                    int64_t n_ray, const float* M_inv, const float* b, float* out,
                                                    ctlab_stream_t stream) {
    static const char* fn = "ctlab_forward_project_3d_table_ranged";
    const int64_t n[3] = {n_x, n_y, n_z};
    int st = check_table_geometry(fn, n, TableArgs{tvals, u16, start, scale, K}, src, dst, n_ray, M_inv, b);
    if (st) return st;
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range [1, 65535]", fn, (long long)batch);
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(volume && out && first && end && (!u16 || k_first), "%s: null pointer", fn);
    Dims3 dims;
    for (int a = 0; a < 3; ++a) dims.n[a] = (int)n[a];
    const int64_t V = n_x * n_y * n_z;
    const dim3 grid(grid_for(n_ray), (unsigned)batch);
    if (u16)
        hipLaunchKernelGGL((k_forward_table_ranged<true>), grid, dim3(kBlock), 0, as_stream(stream), volume, out,
                           n_ray, V, dims, tvals, u16, start, scale, K, first, end, k_first, src, dst, M_inv, b);
    else
        hipLaunchKernelGGL((k_forward_table_ranged<false>), grid, dim3(kBlock), 0, as_stream(stream), volume, out,
                           n_ray, V, dims, tvals, u16, start, scale, K, first, end, k_first, src, dst, M_inv, b);
    return check_launch(fn);
}
