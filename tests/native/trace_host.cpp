// Test-only: runs the product's traversal (siddon_traverse.h) on the CPU so its
// segment lists can be compared with the oracle without a GPU.
#include "../../ct_laboratory_amd/csrc/siddon_traverse.h"

using namespace ctlab;

template <int ND, int MODE, bool DIAG>
static int64_t run(const Grid<ND>& g, const float* src, const float* dst, int64_t n_ray, int64_t cap,
                   int64_t* idx, float* len, int64_t* counts) {
    int64_t worst = 0;
    for (int64_t r = 0; r < n_ray; ++r) {
        int64_t k = 0;
        int64_t* io = idx + r * cap;
        float* lo = len + r * cap;
        trace_ray<ND, MODE, DIAG>(g, src + r * ND, dst + r * ND, [&](int i, float l) {
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
        return g.diag ? run<ND, MODE_TORCH, true>(g, src, dst, n_ray, cap, idx, len, counts)
                      : run<ND, MODE_TORCH, false>(g, src, dst, n_ray, cap, idx, len, counts);
    }
    return g.diag ? run<ND, MODE_OTF, true>(g, src, dst, n_ray, cap, idx, len, counts)
                  : run<ND, MODE_OTF, false>(g, src, dst, n_ray, cap, idx, len, counts);
}

extern "C" int64_t host_trace(int ndim, int mode, int zero_rule, const int* n, const float* minv,
                              const float* b, const float* src, const float* dst, int64_t n_ray,
                              int64_t cap, int64_t* idx, float* len, int64_t* counts) {
    if (ndim == 2) return dispatch<2>(mode, zero_rule, n, minv, b, src, dst, n_ray, cap, idx, len, counts);
    return dispatch<3>(mode, zero_rule, n, minv, b, src, dst, n_ray, cap, idx, len, counts);
}
