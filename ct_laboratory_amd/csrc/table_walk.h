// table_walk.h -- walks along precomputed crossing-table rows (float tables
// and uint16 delta tables), shared by the table projectors (ctlab_tvals.hip)
// and the table brick back projector (ctlab_table_brick.hip).
//
// A walker hands each segment that passes its table's length rule to
// visit(j, k_prev, c, len): j is the index of the segment's end entry (the
// segment runs from entry j-1 to entry j), k_prev the uint16 walk's integer
// running sum at entry j-1 (0 for float tables), c the midpoint's continuous
// voxel coordinates and len the world length.  The visitor does the voxel
// lookup, so every consumer rounds the same coordinates the same way.
#pragma once

#include "ctlab_common.h"
#include "siddon_traverse.h"

namespace ctlab {

// A block's rows are read through LDS: the block stages C entries of each of
// its kBlock rows with coalesced loads (consecutive lanes read consecutive
// entries of one row), then each lane continues its own row's walk over them.
// A lane sees its row's entries in order, so its visits, and any sum it forms,
// are those of a plain walk along the row.  Lanes without a row (past n_ray,
// or with nothing to add) still stage, so every lane reaches every barrier.
template <typename T, int C>
struct RowStage {
    static constexpr int kStride = C + (sizeof(T) == 4 ? 1 : 2);  // rows start on different banks
    T e[kBlock * kStride];
};

template <typename T, int C>
__device__ __forceinline__ void stage_rows(RowStage<T, C>& st, const T* __restrict__ table, int64_t K, int64_t r0,
                                           int64_t n_ray, int64_t c0) {
    __syncthreads();  // the previous chunk has been consumed
    for (int e = threadIdx.x; e < kBlock * C; e += kBlock) {
        const int row = e / C, col = e % C;
        const int64_t rr = r0 + row, k = c0 + col;
        if (rr < n_ray && k < K) st.e[row * RowStage<T, C>::kStride + col] = table[rr * K + k];
    }
    __syncthreads();
}

// 64 bytes of each row per chunk (17 KB of LDS per block: 256 rows x 17 floats,
// or 256 rows x 34 uint16 entries).  On a 1.6 GB 3D
// table, 32 and 64 bytes per row ran 1.19 and 1.41 ms, 128 bytes 1.67 ms and
// 256 bytes 2.86 ms: smaller stages leave room for more blocks per CU.
constexpr int kStageF32 = 16;
constexpr int kStageU16 = 32;

// Float table (3d.cu:272-376, 3d_torch.py:219-271): segment j runs from
// entry j-1 to entry j; INF ends a sorted row and a finite value after it
// restarts the walk; segments of length 0 are skipped.
template <int ND, bool DIAG, class Visit>
__device__ __forceinline__ void walk_table(const Grid<ND>& g, const float* s, const float* d,
                                           const float* __restrict__ tvals, int64_t K, int64_t r0, int64_t n_ray,
                                           bool active, RowStage<float, kStageF32>& st, Visit&& visit) {
    float vw[ND], p0[ND];
#pragma unroll
    for (int a = 0; a < ND; ++a) vw[a] = d[a] - s[a];
    if (K < 2 || !g.finite) return;  // uniform over the block
    MidIndex<ND, DIAG> mi;
    mi.init(g);
    float t0 = 0.0f;
    const float* __restrict__ mine = st.e + threadIdx.x * RowStage<float, kStageF32>::kStride;
    for (int64_t c0 = 0; c0 < K; c0 += kStageF32) {
        stage_rows(st, tvals, K, r0, n_ray, c0);
        if (!active) continue;
        const int jn = (int)min((int64_t)kStageF32, K - c0);
        for (int j = 0; j < jn; ++j) {
            const float t1 = mine[j];
            if (c0 + j == 0 || isinf(t0) || isinf(t1)) {
                // Entry 0 starts the pair walk; INF marks the end of a sorted
                // row, and a finite value after it restarts the walk.
                t0 = t1;
#pragma unroll
                for (int a = 0; a < ND; ++a) p0[a] = s[a] + t0 * vw[a];
                continue;
            }
            float c[ND];
            const float len = segment_mid<ND, DIAG>(mi, s, vw, p0, t1, c);
            t0 = t1;
            if (len > 0.0f) visit(c0 + j, 0u, c, len);
        }
    }
}

// uint16 table with the CUDA decode (3d.cu:99-157, 215-269): entry 0 holds no
// delta and t(0) = start; t(j) = start + k(j) * scale in one fused multiply-add
// with the integer running sum k(j); segments shorter than 1e-9 are skipped
// (3d.cu:137).
template <bool DIAG, class Visit>
__device__ __forceinline__ void walk_u16(const Grid<3>& g, const float* s, const float* d,
                                         const uint16_t* __restrict__ u16, int64_t K, int64_t r0, int64_t n_ray,
                                         bool active, float st0, float sc, RowStage<uint16_t, kStageU16>& st,
                                         Visit&& visit) {
    float vw[3], p0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        vw[a] = d[a] - s[a];
        p0[a] = s[a] + st0 * vw[a];
    }
    if (!g.finite || K < 2) return;  // uniform over the block
    MidIndex<3, DIAG> mi;
    mi.init(g);
    uint32_t ks = 0;
    const uint16_t* __restrict__ mine = st.e + threadIdx.x * RowStage<uint16_t, kStageU16>::kStride;
    for (int64_t c0 = 0; c0 < K; c0 += kStageU16) {
        stage_rows(st, u16, K, r0, n_ray, c0);
        if (!active) continue;
        const int jn = (int)min((int64_t)kStageU16, K - c0);
        for (int j = (c0 == 0) ? 1 : 0; j < jn; ++j) {  // entry 0 holds no delta
            const uint32_t dk = (uint32_t)mine[j];
            // A zero delta repeats the previous t bit for bit: its segment has length
            // exactly 0 (skipped below) and leaves P(t0) as it is.  The padding of a
            // row (INF -> delta 0) is all such entries.
            if (dk == 0u) continue;
            const uint32_t kp = ks;
            ks += dk;
            // t1 = start + k * scale with the integer running sum k (3d.cu:128-129)
            const float t1 = __builtin_fmaf((float)ks, sc, st0);
            float c[3];
            const float len = segment_mid<3, DIAG>(mi, s, vw, p0, t1, c);
            if (!(len >= 1e-9f)) continue;  // 3d.cu:137
            visit(c0 + j, kp, c, len);
        }
    }
}

}  // namespace ctlab
