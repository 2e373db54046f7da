// Test-only: runs the product's traversal (siddon_traverse.h) on the CPU so its
// segment lists can be compared with the oracle without a GPU.
#include <string.h>

#include "../../ct_laboratory_amd/csrc/siddon_traverse.h"

using namespace ctlab;

template <int ND, int MODE, bool DIAG, bool POW2 = false>
static int64_t run(const Grid<ND>& g, const float* src, const float* dst, int64_t n_ray, int64_t cap,
                   int64_t* idx, float* len, int64_t* counts) {
    int64_t worst = 0;
    for (int64_t r = 0; r < n_ray; ++r) {
        int64_t k = 0;
        int64_t* io = idx + r * cap;
        float* lo = len + r * cap;
        trace_ray<ND, MODE, DIAG, POW2>(g, src + r * ND, dst + r * ND, [&](int i, float l) {
            if (k < cap) { io[k] = i; lo[k] = l; }
            ++k;
        });
        counts[r] = k;
        if (k > worst) worst = k;
    }
    return worst;
}

template <int ND>
static int64_t dispatch(int mode, int zero_rule, const int* n, const float* minv, const float* b,
                        const float* src, const float* dst, int64_t n_ray, int64_t cap, int64_t* idx,
                        float* len, int64_t* counts) {
    Grid<ND> g = load_grid<ND>(n, minv, b, zero_rule);
    if (mode == MODE_TORCH) {
        if (g.pow2) return run<ND, MODE_TORCH, true, true>(g, src, dst, n_ray, cap, idx, len, counts);
        return g.diag ? run<ND, MODE_TORCH, true>(g, src, dst, n_ray, cap, idx, len, counts)
                      : run<ND, MODE_TORCH, false>(g, src, dst, n_ray, cap, idx, len, counts);
    }
    return g.diag ? run<ND, MODE_OTF, true>(g, src, dst, n_ray, cap, idx, len, counts)
                  : run<ND, MODE_OTF, false>(g, src, dst, n_ray, cap, idx, len, counts);
}

// The reference's 2D on-the-fly march (MODE_MARCH): every in-grid step.
static int64_t run_march(const Grid<2>& g, const float* src, const float* dst, int64_t n_ray, int64_t cap,
                         int64_t* idx, float* len, int64_t* counts) {
    int64_t worst = 0;
    for (int64_t r = 0; r < n_ray; ++r) {
        int64_t k = 0;
        int64_t* io = idx + r * cap;
        float* lo = len + r * cap;
        march_ray_2d(g, src + r * 2, dst + r * 2, [&](uint32_t i, float l) {
            if (k < cap) { io[k] = i; lo[k] = l; }
            ++k;
        });
        counts[r] = k;
        if (k > worst) worst = k;
    }
    return worst;
}

extern "C" int64_t host_trace(int ndim, int mode, int zero_rule, const int* n, const float* minv,
                              const float* b, const float* src, const float* dst, int64_t n_ray,
                              int64_t cap, int64_t* idx, float* len, int64_t* counts) {
    if (ndim == 2 && mode == MODE_MARCH)
        return run_march(load_grid<2>(n, minv, b, zero_rule), src, dst, n_ray, cap, idx, len, counts);
    if (ndim == 2) return dispatch<2>(mode, zero_rule, n, minv, b, src, dst, n_ray, cap, idx, len, counts);
    return dispatch<3>(mode, zero_rule, n, minv, b, src, dst, n_ray, cap, idx, len, counts);
}

// Window variant: segments whose voxel lies in [lo, lo+ext), reported with
// their full-grid flat index.
template <int ND, int MODE, bool DIAG, bool POW2 = false>
static void run_window(const Grid<ND>& g, const Window<ND>& w, const float* src, const float* dst, int64_t n_ray,
                       int64_t cap, int64_t* idx, float* len, int64_t* counts) {
    for (int64_t r = 0; r < n_ray; ++r) {
        int64_t k = 0;
        int64_t* io = idx + r * cap;
        float* lo = len + r * cap;
        trace_ray_window<ND, MODE, DIAG, true, POW2>(g, w, src + r * ND, dst + r * ND, [&](uint32_t li, float l) {
            // window index (compact strides) -> grid flat index
            int64_t rem = li, gi = 0;
            int64_t loc[ND];
            for (int a = 0; a < ND; ++a) {
                loc[a] = rem / w.stride[a];
                rem %= w.stride[a];
            }
            for (int a = 0; a < ND; ++a) gi = gi * g.n[a] + (loc[a] + w.lo[a]);
            if (k < cap) { io[k] = gi; lo[k] = l; }
            ++k;
        });
        counts[r] = k;
    }
}

extern "C" int64_t host_trace_window(int ndim, int mode, int zero_rule, const int* n, const float* minv,
                                     const float* b, const int* lo, const int* ext, const float* src,
                                     const float* dst, int64_t n_ray, int64_t cap, int64_t* idx, float* len,
                                     int64_t* counts) {
    if (ndim == 3) {
        Grid<3> g = load_grid<3>(n, minv, b, zero_rule);
        Window<3> w;
        for (int a = 0; a < 3; ++a) { w.lo[a] = lo[a]; w.ext[a] = ext[a]; }
        w.stride[2] = 1; w.stride[1] = ext[2]; w.stride[0] = ext[1] * ext[2];
        if (mode == MODE_TORCH) {
            if (g.pow2) run_window<3, MODE_TORCH, true, true>(g, w, src, dst, n_ray, cap, idx, len, counts);
            else if (g.diag) run_window<3, MODE_TORCH, true>(g, w, src, dst, n_ray, cap, idx, len, counts);
            else run_window<3, MODE_TORCH, false>(g, w, src, dst, n_ray, cap, idx, len, counts);
        } else {
            if (g.diag) run_window<3, MODE_OTF, true>(g, w, src, dst, n_ray, cap, idx, len, counts);
            else run_window<3, MODE_OTF, false>(g, w, src, dst, n_ray, cap, idx, len, counts);
        }
    } else {
        Grid<2> g = load_grid<2>(n, minv, b, zero_rule);
        Window<2> w;
        for (int a = 0; a < 2; ++a) { w.lo[a] = lo[a]; w.ext[a] = ext[a]; }
        w.stride[1] = 1; w.stride[0] = ext[1];
        if (mode == MODE_TORCH) {
            if (g.pow2) run_window<2, MODE_TORCH, true, true>(g, w, src, dst, n_ray, cap, idx, len, counts);
            else if (g.diag) run_window<2, MODE_TORCH, true>(g, w, src, dst, n_ray, cap, idx, len, counts);
            else run_window<2, MODE_TORCH, false>(g, w, src, dst, n_ray, cap, idx, len, counts);
        } else {
            if (g.diag) run_window<2, MODE_OTF, true>(g, w, src, dst, n_ray, cap, idx, len, counts);
            else run_window<2, MODE_OTF, false>(g, w, src, dst, n_ray, cap, idx, len, counts);
        }
    }
    return 0;
}

// Every float bit pattern in [lo, hi): counts where half_exact / twice_exact
// differ in bits from the host's 0.5f * x and 2.0f * x (NaNs compare as NaN).
extern "C" int64_t host_scale_mismatches(uint64_t lo, uint64_t hi) {
    int64_t bad = 0;
    for (uint64_t k = lo; k < hi; ++k) {
        const uint32_t u = (uint32_t)k;
        float x;
        ::memcpy(&x, &u, 4);
        const float h = ctlab::half_exact(x), t = ctlab::twice_exact(x);
        volatile float hx = 0.5f * x, tx = 2.0f * x;
        const float hr = hx, tr = tx;
        uint32_t a, b, c, d;
        ::memcpy(&a, &h, 4);
        ::memcpy(&b, &hr, 4);
        ::memcpy(&c, &t, 4);
        ::memcpy(&d, &tr, 4);
        if (x != x) {
            bad += !(h != h) + !(t != t);
        } else {
            bad += (a != b) + (c != d);
        }
    }
    return bad;
}
