// ctlab_batch.hip -- many 2D scans of different geometry and size in one
// persistent launch.
//
// Each scan is described by a ctlab_scan2d record: its image, sinogram, rays,
// affine map, segment rule and the index of its first work item.  A work item
// is a run of rays_per_item consecutive rays of one scan.  The launch has a
// fixed number of workgroups; each one takes the next item from a global
// counter (one 64-bit atomic per item) and traces the item's rays, one ray per
// lane, with the same traversal as the single-scan kernels.  Small scans thus
// share the machine instead of each under-filling it in its own launch, and
// uneven scans balance through the queue.  Items of one scan are contiguous in
// the queue, so concurrently running workgroups mostly work on one image.
#include "ctlab_common.h"
#include "siddon_traverse.h"

namespace ctlab {

static_assert(sizeof(ctlab_scan2d) == 88, "ctlab_scan2d layout");

__device__ __forceinline__ int find_scan(const ctlab_scan2d* __restrict__ scans, int n_scan, int64_t item) {
    int lo = 0, hi = n_scan;  // largest k with item_begin <= item
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (scans[mid].item_begin <= item) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ Grid<2> scan_grid(const ctlab_scan2d& sc) {
    const int n[2] = {sc.n_row, sc.n_col};
    return load_grid<2>(n, sc.minv, sc.b, sc.zero_rule);
}

template <int MODE, bool DIAG, bool BACK>
__device__ __forceinline__ void item_rays(const ctlab_scan2d& sc, const Grid<2>& g, int64_t r0, int64_t r1) {
    const float* __restrict__ src = reinterpret_cast<const float*>(sc.src);
    const float* __restrict__ dst = reinterpret_cast<const float*>(sc.dst);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const float s[2] = {src[2 * r], src[2 * r + 1]};
        const float d[2] = {dst[2 * r], dst[2 * r + 1]};
        if (BACK) {
            const float val = reinterpret_cast<const float*>(sc.sino)[r];
            if (val == 0.0f) continue;
            float* __restrict__ img = reinterpret_cast<float*>(sc.image);
            trace_ray<2, MODE, DIAG>(g, s, d, [&](uint32_t idx, float len) { unsafeAtomicAdd(img + idx, val * len); });
        } else {
            const float* __restrict__ img = reinterpret_cast<const float*>(sc.image);
            float acc = 0.0f, pv = 0.0f, pl = 0.0f;  // gather consumed one segment later
            trace_ray<2, MODE, DIAG>(g, s, d, [&](uint32_t idx, float len) {
                acc = __builtin_fmaf(pv, pl, acc);
                pv = img[idx];
                pl = len;
            });
            reinterpret_cast<float*>(sc.sino)[r] = __builtin_fmaf(pv, pl, acc);
        }
    }
}

template <bool BACK>
__global__ __launch_bounds__(kBlock) void k_batch2d(const ctlab_scan2d* __restrict__ scans, int n_scan,
                                                    int64_t total_items, int64_t rays_per_item,
                                                    unsigned long long* __restrict__ counter) {
    __shared__ int64_t s_item;
    while (true) {
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t item = s_item;
        __syncthreads();  // every lane has read s_item before it is overwritten
        if (item >= total_items) break;  // uniform: every workgroup reaches this exit
        const ctlab_scan2d sc = scans[find_scan(scans, n_scan, item)];
        const Grid<2> g = scan_grid(sc);
        const int64_t r0 = (item - sc.item_begin) * rays_per_item;
        const int64_t r1 = min(r0 + rays_per_item, sc.n_ray);
        if (!g.finite) continue;
        if (sc.mode == MODE_OTF) {
            if (g.diag) item_rays<MODE_OTF, true, BACK>(sc, g, r0, r1);
            else item_rays<MODE_OTF, false, BACK>(sc, g, r0, r1);
        } else {
            if (g.diag) item_rays<MODE_TORCH, true, BACK>(sc, g, r0, r1);
            else item_rays<MODE_TORCH, false, BACK>(sc, g, r0, r1);
        }
    }
}

// Zeroes every scan's image (grid-stride over the scans, then their pixels) and
// the item counter.
__global__ __launch_bounds__(kBlock) void k_batch2d_clear(const ctlab_scan2d* __restrict__ scans, int n_scan,
                                                          int clear_images, unsigned long long* __restrict__ counter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0ull;
    if (!clear_images) return;
    for (int k = blockIdx.y; k < n_scan; k += gridDim.y) {
        float* __restrict__ img = reinterpret_cast<float*>(scans[k].image);
        const int64_t n = (int64_t)scans[k].n_row * scans[k].n_col;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
            img[i] = 0.0f;
    }
}

int batch2d_clear(const char* fn, const ctlab_scan2d* scans, int64_t n_scan, bool clear_images,
                  unsigned long long* counter, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch2d_clear, dim3(64, (unsigned)(n_scan < 1024 ? n_scan : 1024)), dim3(kBlock), 0, stream,
                       scans, (int)n_scan, clear_images ? 1 : 0, counter);
    return check_launch(fn);
}

static int batch2d(const char* fn, const ctlab_scan2d* scans, int64_t n_scan, int64_t total_items,
                   int64_t rays_per_item, int64_t n_groups, void* counter, bool back, int accumulate,
                   hipStream_t stream) {
    CTLAB_REQUIRE(n_scan >= 0 && n_scan <= 0x7fffffffLL, "%s: n_scan out of range", fn);
    CTLAB_REQUIRE(total_items >= 0 && rays_per_item >= 1 && rays_per_item <= (1LL << 30), "%s: bad work items", fn);
    CTLAB_REQUIRE(n_groups >= 1 && n_groups <= 65536, "%s: n_groups %lld out of range", fn, (long long)n_groups);
    if (n_scan == 0) return CTLAB_OK;
    CTLAB_REQUIRE(scans && counter, "%s: null pointer", fn);
    auto* ctr = reinterpret_cast<unsigned long long*>(counter);
    int st = batch2d_clear(fn, scans, n_scan, back && !accumulate, ctr, stream);
    if (st) return st;
    if (total_items == 0) return CTLAB_OK;
    if (back)
        hipLaunchKernelGGL(k_batch2d<true>, dim3((unsigned)n_groups), dim3(kBlock), 0, stream, scans, (int)n_scan,
                           total_items, rays_per_item, ctr);
    else
        hipLaunchKernelGGL(k_batch2d<false>, dim3((unsigned)n_groups), dim3(kBlock), 0, stream, scans, (int)n_scan,
                           total_items, rays_per_item, ctr);
    return check_launch(fn);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int ctlab_forward_project_2d_batch(const ctlab_scan2d* scans, int64_t n_scan, int64_t total_items,
                                              int64_t rays_per_item, int64_t n_groups, void* counter,
                                              ctlab_stream_t stream) {
    return batch2d("forward_project_2d_batch", scans, n_scan, total_items, rays_per_item, n_groups, counter, false,
                   0, as_stream(stream));
}

extern "C" int ctlab_back_project_2d_batch(const ctlab_scan2d* scans, int64_t n_scan, int64_t total_items,
                                           int64_t rays_per_item, int64_t n_groups, void* counter, int accumulate,
                                           ctlab_stream_t stream) {
    return batch2d("back_project_2d_batch", scans, n_scan, total_items, rays_per_item, n_groups, counter, true,
                   accumulate, as_stream(stream));
}
