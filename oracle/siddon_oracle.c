/*
 * siddon_oracle.c -- CPU restatement of ct_laboratory's Siddon projector semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and never by the product library (ct_laboratory_amd/).  It is deliberately the
 * slow, sort-based algorithm of the reference (enumerate every plane crossing,
 * sort, walk consecutive pairs), so that it is independent of the merge traversal
 * the HIP kernels use.
 *
 * Built with -O2 -ffp-contract=off: every float multiply and add is rounded on
 * its own, as in the reference's torch path (each torch op is a separate kernel).
 * Where the reference's CUDA code would be contracted by nvcc into a fused
 * multiply-add (nvcc --fmad=true default), this file writes fmaf() explicitly.
 *
 * Semantics restated (citations are reference files):
 *   MODE_TORCH  ct_laboratory/tomography/ct_projector_3d_torch.py:78-158 (crossings,
 *               division, |dir|<1e-12 -> INF, t outside [0,1] -> INF, sort) and
 *               :161-386 (world points src+t*(dst-src), seg = sqrt(dx^2+dy^2+dz^2),
 *               midpoint 0.5*(p0+p1), index = round-half-even(M_inv (mid-b)),
 *               out-of-range segments dropped).  2D: ct_projector_2d_torch.py:4-250
 *               with the exact-zero rule dir==0 (:42).
 *   MODE_OTF    src/ct_projector_3d.cu:834-1012 (ray_kernel_plane_sort): crossings
 *               (plane-s)*(1/dir) for |dir|>1e-12, kept if 0<=t<=1, endpoints 0 and 1
 *               inserted, sorted, pairs with t1<=t0 skipped, voxel from the world
 *               point at t_mid, seg = (t1-t0)*|dst-src|.
 *   MODE_MARCH  src/ct_projector_2d.cu:246-331 and 334-413 (2D on-the-fly kernels):
 *               slab entry/exit, entry voxel by floor (ceil - 1 going down) of the
 *               entry coordinate, incremental march with t_next += |1/v|, every
 *               in-grid step visited with length (t_hit - t_cur)*|dst-src|.
 *   F64         MODE_TORCH evaluated in double precision (adjoint and closed-form
 *               reference).
 *
 * Parity pinning of this file: see tests/test_oracle.py (closed forms from the
 * reference's python_examples, brute-force FP64 enumeration, adjoint identity,
 * chord length A.1, golden vectors written from the reference's torch functions).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MODE_TORCH 0
#define MODE_OTF 1
#define MODE_MARCH 2

#define ZERO_RULE_THRESH 0 /* |dir| < 1e-12 gives no crossings (3D torch, all CUDA) */
#define ZERO_RULE_EXACT 1  /* dir == 0 gives no crossings (2D torch, 2d_torch.py:42) */

typedef struct {
    int ndim;
    int64_t n[3];
    float minv[9]; /* row-major ndim x ndim inverse of M */
    float b[3];
    int zero_rule;
} orc_geom;

static int cmp_f32(const void* a, const void* b) {
    float x = *(const float*)a, y = *(const float*)b;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}
static int cmp_f64(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}

/* Voxel coordinate of a segment midpoint for the index rule (3d_torch.py:244-254;
 * 2d_torch.py:120-129): elementwise tensor operations, M_inv[a][j]*shift_j
 * summed left to right with separate rounding. */
static void index_coords_f32(const orc_geom* g, const float* m, float* out) {
    int nd = g->ndim;
    float sh[3];
    for (int j = 0; j < nd; ++j) sh[j] = m[j] - g->b[j];
    for (int a = 0; a < nd; ++a) {
        float acc = g->minv[a * nd + 0] * sh[0];
        for (int j = 1; j < nd; ++j) acc = acc + g->minv[a * nd + j] * sh[j];
        out[a] = acc;
    }
}

/* Voxel coordinates of a ray end point.  The torch projector forms them as a
 * matrix product, (p - b) @ M_inv^T (3d_torch.py:100-103; 2d_torch.py:26-29);
 * torch's CPU float32 matmul evaluates each 2- or 3-term dot product as the
 * chain fma(x2, m2, fma(x1, m1, x0*m0)) (checked bit for bit against torch, and
 * pinned by the golden vectors of tests/golden).  The reference's CUDA
 * apply_affine_inverse_3d (3d.cu:21-36) contracts to the same chain under
 * nvcc.  For a diagonal M_inv every form gives (p_a - b_a) * M_inv[a][a]. */
static void to_voxel_mm(const orc_geom* g, const float* p, float* out) {
    int nd = g->ndim;
    float sh[3];
    for (int j = 0; j < nd; ++j) sh[j] = p[j] - g->b[j];
    for (int a = 0; a < nd; ++a) {
        float acc = g->minv[a * nd + 0] * sh[0];
        for (int j = 1; j < nd; ++j) acc = fmaf(g->minv[a * nd + j], sh[j], acc);
        out[a] = acc;
    }
}

static int64_t flat_index(const orc_geom* g, const int64_t* ii) {
    if (g->ndim == 2) return ii[0] * g->n[1] + ii[1];
    return (ii[0] * g->n[1] + ii[1]) * g->n[2] + ii[2];
}

/* Round half to even and bounds-check; returns 0 if out of range. */
static int rne_in_range_f32(const orc_geom* g, const float* c, int64_t* ii) {
    for (int a = 0; a < g->ndim; ++a) {
        float r = nearbyintf(c[a]); /* default FE_TONEAREST: ties to even */
        if (!(r >= 0.0f) || !(r < (float)g->n[a])) return 0;
        ii[a] = (int64_t)r;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Sorted crossings of one ray                                                */
/* ------------------------------------------------------------------------ */

/* MODE_TORCH list: all n_a+1 planes per axis, INF when excluded, sorted.
 * Returns list length (always sum(n_a+1)). */
static int64_t crossings_torch_f32(const orc_geom* g, const float* src, const float* dst, float* t) {
    float s[3], d[3], dir[3];
    to_voxel_mm(g, src, s);
    to_voxel_mm(g, dst, d);
    int64_t cnt = 0;
    for (int a = 0; a < g->ndim; ++a) {
        dir[a] = d[a] - s[a];
        int zero = (g->zero_rule == ZERO_RULE_EXACT) ? (dir[a] == 0.0f) : (fabsf(dir[a]) < 1.0e-12f);
        for (int64_t p = 0; p <= g->n[a]; ++p) {
            float plane = (float)p - 0.5f;
            float tt = (plane - s[a]) / dir[a];
            if (zero) tt = INFINITY;
            if (tt < 0.0f || tt > 1.0f) tt = INFINITY;
            t[cnt++] = tt;
        }
    }
    qsort(t, (size_t)cnt, sizeof(float), cmp_f32);
    return cnt;
}

/* MODE_OTF list (3d.cu:852-895): kept crossings, plus 0 and 1, sorted. */
static int64_t crossings_otf_f32(const orc_geom* g, const float* src, const float* dst, float* t) {
    float s[3], d[3];
    to_voxel_mm(g, src, s);
    to_voxel_mm(g, dst, d);
    int64_t cnt = 0;
    for (int a = 0; a < g->ndim; ++a) {
        float v = d[a] - s[a];
        if (fabsf(v) > 1.0e-12f) {
            float inv = 1.0f / v;
            for (int64_t p = 0; p <= g->n[a]; ++p) {
                float tt = ((float)p - 0.5f - s[a]) * inv;
                if (tt >= 0.0f && tt <= 1.0f) t[cnt++] = tt;
            }
        }
    }
    t[cnt++] = 0.0f;
    t[cnt++] = 1.0f;
    qsort(t, (size_t)cnt, sizeof(float), cmp_f32);
    return cnt;
}

static float ray_len_otf(const orc_geom* g, const float* src, const float* dst) {
    float v[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < g->ndim; ++a) v[a] = dst[a] - src[a];
    float acc = v[0] * v[0];
    for (int a = 1; a < g->ndim; ++a) acc = fmaf(v[a], v[a], acc);
    return sqrtf(acc);
}

/* (int) of a float with the GPU conversion's behaviour for NaN (0) and large
 * magnitudes (saturated; here at +-2^30, out of every grid either way). */
static int cvt_i32(float x) {
    if (isnan(x)) return 0;
    if (x >= 1073741824.0f) return 1073741824;
    if (x <= -1073741824.0f) return -1073741824;
    return (int)x;
}

/* MODE_MARCH (2d.cu:246-331): the steps of the reference's 2D on-the-fly
 * march.  nvcc contracts a*b + c by default, so those forms are fmaf here: the
 * affine map (2d.cu:25-28), the entry point (:289-290) and the first crossing
 * (:295-300); the accumulation t_next += dt is a plain add.  The ray length is
 * the double sum of squares, square-rooted and rounded to float (the kernel
 * calls hypotf).  Every step in the grid is written, whatever its length. */
static int64_t trace_march2d_f32(const orc_geom* g, const float* src, const float* dst, int64_t* idx_out,
                                 float* len_out) {
    float vx = dst[0] - src[0], vy = dst[1] - src[1];
    float ray_len = (float)sqrt(fma((double)vx, (double)vx, (double)vy * (double)vy));
    if (ray_len < 1e-12f) return 0;
    float s[2], d[2];
    to_voxel_mm(g, src, s);
    to_voxel_mm(g, dst, d);
    float sr = s[0], sc = s[1];
    float vr = d[0] - sr, vc = d[1] - sc;
    float inv_vr = (fabsf(vr) < 1e-12f) ? 1e32f : 1.0f / vr;
    float inv_vc = (fabsf(vc) < 1e-12f) ? 1e32f : 1.0f / vc;
    int n_row = (int)g->n[0], n_col = (int)g->n[1];
    float t_r0 = (-0.5f - sr) * inv_vr;
    float t_r1 = (((float)n_row - 0.5f) - sr) * inv_vr;
    float t_c0 = (-0.5f - sc) * inv_vc;
    float t_c1 = (((float)n_col - 0.5f) - sc) * inv_vc;
    float t_enter = fmaxf(fminf(t_r0, t_r1), fminf(t_c0, t_c1));
    float t_exit = fminf(fmaxf(t_r0, t_r1), fmaxf(t_c0, t_c1));
    if (t_exit <= t_enter) return 0;
    t_enter = fmaxf(t_enter, 0.0f);
    t_exit = fminf(t_exit, 1.0f);
    float r_enter = fmaf(t_enter, vr, sr);
    float c_enter = fmaf(t_enter, vc, sc);
    int ri = (vr >= 0.0f) ? cvt_i32(floorf(r_enter)) : cvt_i32(ceilf(r_enter)) - 1;
    int ci = (vc >= 0.0f) ? cvt_i32(floorf(c_enter)) : cvt_i32(ceilf(c_enter)) - 1;
    float t_next_r = (vr >= 0.0f) ? fmaf(((float)ri + 0.5f) - r_enter, inv_vr, t_enter)
                                  : fmaf(((float)ri - 0.5f) - r_enter, inv_vr, t_enter);
    float t_next_c = (vc >= 0.0f) ? fmaf(((float)ci + 0.5f) - c_enter, inv_vc, t_enter)
                                  : fmaf(((float)ci - 0.5f) - c_enter, inv_vc, t_enter);
    float dt_r = fabsf(inv_vr), dt_c = fabsf(inv_vc);
    float t_cur = t_enter;
    int64_t nseg = 0;
    /* bounded like the kernels' loop; finite rays end long before */
    for (int k = 0; k < 2 * (n_row + n_col) + 8 && t_cur < t_exit; ++k) {
        int step_row = (t_next_r < t_next_c);
        float t_hit = step_row ? t_next_r : t_next_c;
        t_hit = fminf(t_hit, t_exit);
        if (ri >= 0 && ri < n_row && ci >= 0 && ci < n_col) {
            idx_out[nseg] = (int64_t)ri * n_col + ci;
            len_out[nseg] = (t_hit - t_cur) * ray_len;
            ++nseg;
        }
        t_cur = t_hit;
        if (step_row) {
            t_next_r += dt_r;
            ri += (vr >= 0.0f) ? 1 : -1;
        } else {
            t_next_c += dt_c;
            ci += (vc >= 0.0f) ? 1 : -1;
        }
    }
    return nseg;
}

/*
 * Segments of one ray under a float mode.  Writes every in-range segment
 * (flat index, length), including zero-length ones, in traversal order;
 * returns their number.  tbuf must hold sum(n_a+1)+2 floats.
 */
static int64_t trace_f32(const orc_geom* g, int mode, const float* src, const float* dst,
                         float* tbuf, int64_t* idx_out, float* len_out) {
    int nd = g->ndim;
    int64_t nseg = 0;
    if (mode == MODE_MARCH) return (nd == 2) ? trace_march2d_f32(g, src, dst, idx_out, len_out) : 0;
    if (mode == MODE_TORCH) {
        int64_t cnt = crossings_torch_f32(g, src, dst, tbuf);
        float vw[3];
        for (int a = 0; a < nd; ++a) vw[a] = dst[a] - src[a];
        for (int64_t i = 0; i + 1 < cnt; ++i) {
            float t0 = tbuf[i], t1 = tbuf[i + 1];
            if (isinf(t0) || isinf(t1)) continue;
            float p0[3], p1[3], m[3], c[3];
            for (int a = 0; a < nd; ++a) {
                p0[a] = src[a] + t0 * vw[a];
                p1[a] = src[a] + t1 * vw[a];
            }
            float ss = 0.0f;
            for (int a = 0; a < nd; ++a) {
                float dd = p1[a] - p0[a];
                float sq = dd * dd;
                ss = (a == 0) ? sq : ss + sq;
            }
            float len = sqrtf(ss);
            for (int a = 0; a < nd; ++a) m[a] = 0.5f * (p0[a] + p1[a]);
            index_coords_f32(g, m, c);
            int64_t ii[3];
            if (!rne_in_range_f32(g, c, ii)) continue;
            idx_out[nseg] = flat_index(g, ii);
            len_out[nseg] = len;
            ++nseg;
        }
    } else {
        float rl = ray_len_otf(g, src, dst);
        if (rl < 1.0e-12f) return 0;
        int64_t cnt = crossings_otf_f32(g, src, dst, tbuf);
        float vw[3];
        for (int a = 0; a < nd; ++a) vw[a] = dst[a] - src[a];
        for (int64_t i = 0; i + 1 < cnt; ++i) {
            float t0 = tbuf[i], t1 = tbuf[i + 1];
            if (t1 <= t0) continue;
            float tm = 0.5f * (t0 + t1);
            float m[3], c[3];
            for (int a = 0; a < nd; ++a) m[a] = fmaf(tm, vw[a], src[a]);
            to_voxel_mm(g, m, c);
            int64_t ii[3];
            if (!rne_in_range_f32(g, c, ii)) continue;
            idx_out[nseg] = flat_index(g, ii);
            len_out[nseg] = (t1 - t0) * rl;
            ++nseg;
        }
    }
    return nseg;
}

/* FP64 version of MODE_TORCH. */
static int64_t trace_f64(const orc_geom* g, const double* minv, const double* bb, const double* src,
                         const double* dst, double* tbuf, int64_t* idx_out, double* len_out) {
    int nd = g->ndim;
    double s[3], d[3], dir[3], vw[3];
    for (int a = 0; a < nd; ++a) {
        double as = 0.0, ad = 0.0;
        for (int j = 0; j < nd; ++j) {
            as += (src[j] - bb[j]) * minv[a * nd + j];
            ad += (dst[j] - bb[j]) * minv[a * nd + j];
        }
        s[a] = as;
        d[a] = ad;
        dir[a] = d[a] - s[a];
        vw[a] = dst[a] - src[a];
    }
    int64_t cnt = 0;
    for (int a = 0; a < nd; ++a) {
        int zero = (g->zero_rule == ZERO_RULE_EXACT) ? (dir[a] == 0.0) : (fabs(dir[a]) < 1.0e-12);
        for (int64_t p = 0; p <= g->n[a]; ++p) {
            double tt = ((double)p - 0.5 - s[a]) / dir[a];
            if (zero || tt < 0.0 || tt > 1.0) tt = INFINITY;
            tbuf[cnt++] = tt;
        }
    }
    qsort(tbuf, (size_t)cnt, sizeof(double), cmp_f64);
    int64_t nseg = 0;
    for (int64_t i = 0; i + 1 < cnt; ++i) {
        double t0 = tbuf[i], t1 = tbuf[i + 1];
        if (isinf(t0) || isinf(t1)) continue;
        double ss = 0.0, m[3], c[3];
        for (int a = 0; a < nd; ++a) {
            double p0 = src[a] + t0 * vw[a], p1 = src[a] + t1 * vw[a];
            ss += (p1 - p0) * (p1 - p0);
            m[a] = 0.5 * (p0 + p1);
        }
        int64_t ii[3];
        int ok = 1;
        for (int a = 0; a < nd; ++a) {
            double acc = 0.0;
            for (int j = 0; j < nd; ++j) acc += minv[a * nd + j] * (m[j] - bb[j]);
            c[a] = nearbyint(acc);
            if (!(c[a] >= 0.0) || !(c[a] < (double)g->n[a])) { ok = 0; break; }
            ii[a] = (int64_t)c[a];
        }
        if (!ok) continue;
        idx_out[nseg] = flat_index(g, ii);
        len_out[nseg] = sqrt(ss);
        ++nseg;
    }
    return nseg;
}

static int64_t list_capacity(const orc_geom* g) {
    int64_t c = 2;
    for (int a = 0; a < g->ndim; ++a) c += g->n[a] + 1;
    return c;
}

static void make_geom(orc_geom* g, int ndim, const int64_t* n, const float* minv, const float* b, int zero_rule) {
    memset(g, 0, sizeof(*g));
    g->ndim = ndim;
    for (int a = 0; a < ndim; ++a) { g->n[a] = n[a]; g->b[a] = b[a]; }
    for (int i = 0; i < ndim * ndim; ++i) g->minv[i] = minv[i];
    g->zero_rule = zero_rule;
}

static int64_t vol_size(const orc_geom* g) {
    int64_t v = 1;
    for (int a = 0; a < g->ndim; ++a) v *= g->n[a];
    return v;
}

/* ======================================================================== */
/* Exported API (ctypes)                                                      */
/* ======================================================================== */

/* Per-ray segment lists: idx/len are [n_ray, cap] row-major; counts [n_ray]. */
int64_t orc_trace(int ndim, const int64_t* n, const float* minv, const float* b, int zero_rule, int mode,
                  const float* src, const float* dst, int64_t n_ray, int64_t cap, int64_t* idx, float* len,
                  int64_t* counts) {
    orc_geom g;
    make_geom(&g, ndim, n, minv, b, zero_rule);
    int64_t lc = list_capacity(&g);
    int64_t worst = 0;
#pragma omp parallel
    {
        float* tb = (float*)malloc(sizeof(float) * (size_t)lc);
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)lc);
        float* lb = (float*)malloc(sizeof(float) * (size_t)lc);
#pragma omp for schedule(dynamic, 256) reduction(max : worst)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_f32(&g, mode, src + r * ndim, dst + r * ndim, tb, ib, lb);
            int64_t kk = k < cap ? k : cap;
            memcpy(idx + r * cap, ib, sizeof(int64_t) * (size_t)kk);
            memcpy(len + r * cap, lb, sizeof(float) * (size_t)kk);
            counts[r] = k;
            if (k > worst) worst = k;
        }
        free(tb);
        free(ib);
        free(lb);
    }
    return worst;
}

/* Forward projection, float semantics, FP64 accumulation. vol [batch, V]; out [batch, n_ray]. */
int orc_forward(int ndim, const int64_t* n, const float* minv, const float* b, int zero_rule, int mode,
                const float* vol, int64_t batch, const float* src, const float* dst, int64_t n_ray, double* out) {
    orc_geom g;
    make_geom(&g, ndim, n, minv, b, zero_rule);
    int64_t lc = list_capacity(&g), V = vol_size(&g);
#pragma omp parallel
    {
        float* tb = (float*)malloc(sizeof(float) * (size_t)lc);
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)lc);
        float* lb = (float*)malloc(sizeof(float) * (size_t)lc);
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_f32(&g, mode, src + r * ndim, dst + r * ndim, tb, ib, lb);
            for (int64_t bi = 0; bi < batch; ++bi) {
                double acc = 0.0;
                const float* vb = vol + bi * V;
                for (int64_t s = 0; s < k; ++s) acc += (double)(vb[ib[s]] * lb[s]);
                out[bi * n_ray + r] = acc;
            }
        }
        free(tb);
        free(ib);
        free(lb);
    }
    return 0;
}

/* Back projection, float semantics, FP64 accumulation. sino [batch, n_ray]; out [batch, V] (overwritten). */
int orc_back(int ndim, const int64_t* n, const float* minv, const float* b, int zero_rule, int mode,
             const float* sino, int64_t batch, const float* src, const float* dst, int64_t n_ray, double* out) {
    orc_geom g;
    make_geom(&g, ndim, n, minv, b, zero_rule);
    int64_t lc = list_capacity(&g), V = vol_size(&g);
    memset(out, 0, sizeof(double) * (size_t)(V * batch));
#pragma omp parallel
    {
        float* tb = (float*)malloc(sizeof(float) * (size_t)lc);
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)lc);
        float* lb = (float*)malloc(sizeof(float) * (size_t)lc);
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_f32(&g, mode, src + r * ndim, dst + r * ndim, tb, ib, lb);
            for (int64_t bi = 0; bi < batch; ++bi) {
                float sv = sino[bi * n_ray + r];
                double* ob = out + bi * V;
                for (int64_t s = 0; s < k; ++s) {
                    double c = (double)(sv * lb[s]);
#pragma omp atomic
                    ob[ib[s]] += c;
                }
            }
        }
        free(tb);
        free(ib);
        free(lb);
    }
    return 0;
}

/* FP64 forward / back (MODE_TORCH in double). */
int orc_forward_f64(int ndim, const int64_t* n, const double* minv, const double* b, int zero_rule,
                    const double* vol, const double* src, const double* dst, int64_t n_ray, double* out) {
    orc_geom g;
    float fz[9] = {0}, fb[3] = {0};
    make_geom(&g, ndim, n, fz, fb, zero_rule);
    int64_t lc = list_capacity(&g);
#pragma omp parallel
    {
        double* tb = (double*)malloc(sizeof(double) * (size_t)lc);
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)lc);
        double* lb = (double*)malloc(sizeof(double) * (size_t)lc);
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_f64(&g, minv, b, src + r * ndim, dst + r * ndim, tb, ib, lb);
            double acc = 0.0;
            for (int64_t s = 0; s < k; ++s) acc += vol[ib[s]] * lb[s];
            out[r] = acc;
        }
        free(tb);
        free(ib);
        free(lb);
    }
    return 0;
}

int orc_back_f64(int ndim, const int64_t* n, const double* minv, const double* b, int zero_rule,
                 const double* sino, const double* src, const double* dst, int64_t n_ray, double* out) {
    orc_geom g;
    float fz[9] = {0}, fb[3] = {0};
    make_geom(&g, ndim, n, fz, fb, zero_rule);
    int64_t lc = list_capacity(&g), V = vol_size(&g);
    memset(out, 0, sizeof(double) * (size_t)V);
#pragma omp parallel
    {
        double* tb = (double*)malloc(sizeof(double) * (size_t)lc);
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)lc);
        double* lb = (double*)malloc(sizeof(double) * (size_t)lc);
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_f64(&g, minv, b, src + r * ndim, dst + r * ndim, tb, ib, lb);
            for (int64_t s = 0; s < k; ++s) {
                double c = sino[r] * lb[s];
#pragma omp atomic
                out[ib[s]] += c;
            }
        }
        free(tb);
        free(ib);
        free(lb);
    }
    return 0;
}

/* FP64 segment lists (for brute-force comparisons). */
int64_t orc_trace_f64(int ndim, const int64_t* n, const double* minv, const double* b, int zero_rule,
                      const double* src, const double* dst, int64_t n_ray, int64_t cap, int64_t* idx,
                      double* len, int64_t* counts) {
    orc_geom g;
    float fz[9] = {0}, fb[3] = {0};
    make_geom(&g, ndim, n, fz, fb, zero_rule);
    int64_t lc = list_capacity(&g), worst = 0;
    double* tb = (double*)malloc(sizeof(double) * (size_t)lc);
    int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)lc);
    double* lb = (double*)malloc(sizeof(double) * (size_t)lc);
    for (int64_t r = 0; r < n_ray; ++r) {
        int64_t k = trace_f64(&g, minv, b, src + r * ndim, dst + r * ndim, tb, ib, lb);
        int64_t kk = k < cap ? k : cap;
        memcpy(idx + r * cap, ib, sizeof(int64_t) * (size_t)kk);
        memcpy(len + r * cap, lb, sizeof(double) * (size_t)kk);
        counts[r] = k;
        if (k > worst) worst = k;
    }
    free(tb);
    free(ib);
    free(lb);
    return worst;
}

/* compute_intersections_{2d,3d}: sorted crossings [n_ray, sum(n_a+1)] with INF
 * padding (2d_torch.py:4-65, 3d_torch.py:78-158; 2d.cu:36-73, 3d.cu:42-96). */
int orc_intersections(int ndim, const int64_t* n, const float* minv, const float* b, int zero_rule,
                      const float* src, const float* dst, int64_t n_ray, float* out) {
    orc_geom g;
    make_geom(&g, ndim, n, minv, b, zero_rule);
    int64_t n_int = list_capacity(&g) - 2;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < n_ray; ++r) crossings_torch_f32(&g, src + r * ndim, dst + r * ndim, out + r * n_int);
    return 0;
}

/* Forward / back with a given float tvals table [n_ray, K] (2d.cu:76-155,
 * 3d.cu:272-376 consumers; op order of 3d_torch.py:219-271).  Segments with an
 * INF endpoint are skipped; zero-length segments contribute zero. */
static int64_t trace_tvals_f32(const orc_geom* g, const float* src, const float* dst, const float* t,
                               int64_t K, int64_t* idx_out, float* len_out) {
    int nd = g->ndim;
    float vw[3];
    for (int a = 0; a < nd; ++a) vw[a] = dst[a] - src[a];
    int64_t nseg = 0;
    for (int64_t i = 0; i + 1 < K; ++i) {
        float t0 = t[i], t1 = t[i + 1];
        if (isinf(t0) || isinf(t1)) continue;
        float p0[3], p1[3], m[3], c[3], ss = 0.0f;
        for (int a = 0; a < nd; ++a) {
            p0[a] = src[a] + t0 * vw[a];
            p1[a] = src[a] + t1 * vw[a];
            float dd = p1[a] - p0[a];
            float sq = dd * dd;
            ss = (a == 0) ? sq : ss + sq;
            m[a] = 0.5f * (p0[a] + p1[a]);
        }
        index_coords_f32(g, m, c);
        int64_t ii[3];
        if (!rne_in_range_f32(g, c, ii)) continue;
        idx_out[nseg] = flat_index(g, ii);
        len_out[nseg] = sqrtf(ss);
        ++nseg;
    }
    return nseg;
}

int orc_forward_tvals(int ndim, const int64_t* n, const float* minv, const float* b, const float* vol,
                      int64_t batch, const float* tvals, int64_t K, const float* src, const float* dst,
                      int64_t n_ray, double* out) {
    orc_geom g;
    make_geom(&g, ndim, n, minv, b, ZERO_RULE_THRESH);
    int64_t V = vol_size(&g);
#pragma omp parallel
    {
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)(K + 1));
        float* lb = (float*)malloc(sizeof(float) * (size_t)(K + 1));
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_tvals_f32(&g, src + r * ndim, dst + r * ndim, tvals + r * K, K, ib, lb);
            for (int64_t bi = 0; bi < batch; ++bi) {
                double acc = 0.0;
                for (int64_t s = 0; s < k; ++s) acc += (double)(vol[bi * V + ib[s]] * lb[s]);
                out[bi * n_ray + r] = acc;
            }
        }
        free(ib);
        free(lb);
    }
    return 0;
}

int orc_back_tvals(int ndim, const int64_t* n, const float* minv, const float* b, const float* sino,
                   int64_t batch, const float* tvals, int64_t K, const float* src, const float* dst,
                   int64_t n_ray, double* out) {
    orc_geom g;
    make_geom(&g, ndim, n, minv, b, ZERO_RULE_THRESH);
    int64_t V = vol_size(&g);
    memset(out, 0, sizeof(double) * (size_t)(V * batch));
#pragma omp parallel
    {
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)(K + 1));
        float* lb = (float*)malloc(sizeof(float) * (size_t)(K + 1));
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_tvals_f32(&g, src + r * ndim, dst + r * ndim, tvals + r * K, K, ib, lb);
            for (int64_t bi = 0; bi < batch; ++bi) {
                float sv = sino[bi * n_ray + r];
                for (int64_t s = 0; s < k; ++s) {
                    double c = (double)(sv * lb[s]);
#pragma omp atomic
                    out[bi * V + ib[s]] += c;
                }
            }
        }
        free(ib);
        free(lb);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* uint16 compression                                                         */
/* ------------------------------------------------------------------------ */

/* compress_tvals_3d_kernel (3d.cu:160-212): integer running sum, roundf (ties
 * away from zero), INF -> delta 0, clamp [0, 32000]. */
int orc_compress_cuda(const float* tvals, int64_t n_ray, int64_t K, const float* src, const float* dst,
                      double voxel_diag_len, uint16_t* out, float* start, float* scale) {
    float vd = (float)voxel_diag_len;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_ray; ++r) {
        const float* s = src + 3 * r;
        const float* d = dst + 3 * r;
        float vx = d[0] - s[0], vy = d[1] - s[1], vz = d[2] - s[2];
        float rl = sqrtf(fmaf(vz, vz, fmaf(vy, vy, vx * vx)));
        rl = fmaxf(rl, 1e-6f);
        float sc = (vd / 10000.0f) / rl;
        scale[r] = sc;
        const float* t = tvals + r * K;
        uint16_t* o = out + r * K;
        float st = t[0];
        start[r] = st;
        o[0] = 0;
        float inv = 1.0f / sc;
        int64_t kprev = 0;
        for (int64_t i = 1; i < K; ++i) {
            float tv = t[i];
            if (isinf(tv)) { o[i] = 0; continue; }
            float q = roundf((tv - st) * inv);
            /* (unsigned int) conversion as the GPU performs it: saturating */
            int64_t kc = (q < 0.0f) ? 0 : (q >= 4294967296.0f) ? (int64_t)0xffffffffu : (int64_t)(uint32_t)q;
            int64_t diff = (int64_t)(int32_t)kc - (int64_t)(int32_t)kprev;
            if (diff < 0) diff = 0;
            if (diff > 32000) diff = 32000;
            o[i] = (uint16_t)diff;
            kprev += diff;
        }
    }
    return 0;
}

/* compress_tvals_to_uint16 (3d_torch.py:4-65): INF -> 10.0, round half even,
 * float running sum, clamp [0, 32000]. voxel_diag_len is passed in (float32,
 * max norm of M applied to the four body diagonals, :27-35; orc_voxel_diag_len). */
int orc_compress_torch(const float* tvals, int64_t n_ray, int64_t K, const float* src, const float* dst,
                       float voxel_diag_len, int16_t* out, float* start, float* scale) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_ray; ++r) {
        const float* s = src + 3 * r;
        const float* d = dst + 3 * r;
        float vx = d[0] - s[0], vy = d[1] - s[1], vz = d[2] - s[2];
        /* torch.norm(dim=1) of a float32 [n,3] tensor on the CPU: the squares
         * summed as an fma chain, then a correctly rounded sqrtf (checked bit
         * for bit against torch; pinned by the golden vectors) */
        float rl = sqrtf(fmaf(vz, vz, fmaf(vy, vy, vx * vx)));
        if (rl < 1e-6f) rl = 1e-6f;
        float sc = (voxel_diag_len / 10000.0f) / rl;
        scale[r] = sc;
        float inv = 1.0f / sc;
        const float* t = tvals + r * K;
        int16_t* o = out + r * K;
        float st = isinf(t[0]) ? 10.0f : t[0];
        start[r] = st;
        o[0] = 0;
        float kprev = 0.0f;
        for (int64_t i = 1; i < K; ++i) {
            float tv = isinf(t[i]) ? 10.0f : t[i];
            float kc = nearbyintf((tv - st) * inv);
            float df = kc - kprev;
            if (df < 0.0f) df = 0.0f;
            if (df > 32000.0f) df = 32000.0f;
            int16_t di = (int16_t)df;
            o[i] = di;
            kprev += (float)di;
        }
    }
    return 0;
}

/* Decoded crossing of the CUDA consumers (3d.cu:126-129): the integer running
 * sum k of the deltas, then start + k*scale contracted by nvcc into one fma. */
static float decode_cuda(uint32_t k, float sc, float st) { return fmaf((float)k, sc, st); }

/* The decoded table of the CUDA consumers: out[r][0] = start, out[r][i] =
 * decode_cuda(sum of deltas 1..i). */
int orc_decode_u16_cuda(const uint16_t* u, const float* start, const float* scale, int64_t n_ray, int64_t K,
                        float* out) {
    for (int64_t r = 0; r < n_ray; ++r) {
        uint32_t ks = 0;
        if (K > 0) out[r * K] = start[r];
        for (int64_t i = 1; i < K; ++i) {
            ks += (uint32_t)u[r * K + i];
            out[r * K + i] = decode_cuda(ks, scale[r], start[r]);
        }
    }
    return 0;
}

/* Compressed consumers (3d.cu:99-157, 215-269): t1 = start + k*scale (FMA),
 * pairs with seg < 1e-9 skipped; segment geometry in the torch op order. */
static int64_t trace_u16_f32(const orc_geom* g, const float* src, const float* dst, const uint16_t* u,
                             float st, float sc, int64_t K, int64_t* idx_out, float* len_out) {
    float vw[3];
    for (int a = 0; a < 3; ++a) vw[a] = dst[a] - src[a];
    float tprev = st;
    uint32_t ks = 0;
    int64_t nseg = 0;
    for (int64_t i = 0; i + 1 < K; ++i) {
        ks += (uint32_t)u[i + 1];
        float t1 = decode_cuda(ks, sc, st);
        float t0 = tprev;
        tprev = t1;
        float p0[3], p1[3], m[3], c[3], ss = 0.0f;
        for (int a = 0; a < 3; ++a) {
            p0[a] = src[a] + t0 * vw[a];
            p1[a] = src[a] + t1 * vw[a];
            float dd = p1[a] - p0[a];
            float sq = dd * dd;
            ss = (a == 0) ? sq : ss + sq;
            m[a] = 0.5f * (p0[a] + p1[a]);
        }
        float len = sqrtf(ss);
        if (!(len >= 1e-9f)) continue;
        index_coords_f32(g, m, c);
        int64_t ii[3];
        if (!rne_in_range_f32(g, c, ii)) continue;
        idx_out[nseg] = flat_index(g, ii);
        len_out[nseg] = len;
        ++nseg;
    }
    return nseg;
}

int orc_forward_u16(const int64_t* n, const float* minv, const float* b, const float* vol, int64_t batch,
                    const uint16_t* u, const float* start, const float* scale, int64_t K, const float* src,
                    const float* dst, int64_t n_ray, double* out) {
    orc_geom g;
    make_geom(&g, 3, n, minv, b, ZERO_RULE_THRESH);
    int64_t V = vol_size(&g);
#pragma omp parallel
    {
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)(K + 1));
        float* lb = (float*)malloc(sizeof(float) * (size_t)(K + 1));
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_u16_f32(&g, src + 3 * r, dst + 3 * r, u + r * K, start[r], scale[r], K, ib, lb);
            for (int64_t bi = 0; bi < batch; ++bi) {
                double acc = 0.0;
                for (int64_t s = 0; s < k; ++s) acc += (double)(vol[bi * V + ib[s]] * lb[s]);
                out[bi * n_ray + r] = acc;
            }
        }
        free(ib);
        free(lb);
    }
    return 0;
}

int orc_back_u16(const int64_t* n, const float* minv, const float* b, const float* sino, int64_t batch,
                 const uint16_t* u, const float* start, const float* scale, int64_t K, const float* src,
                 const float* dst, int64_t n_ray, double* out) {
    orc_geom g;
    make_geom(&g, 3, n, minv, b, ZERO_RULE_THRESH);
    int64_t V = vol_size(&g);
    memset(out, 0, sizeof(double) * (size_t)(V * batch));
#pragma omp parallel
    {
        int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)(K + 1));
        float* lb = (float*)malloc(sizeof(float) * (size_t)(K + 1));
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < n_ray; ++r) {
            int64_t k = trace_u16_f32(&g, src + 3 * r, dst + 3 * r, u + r * K, start[r], scale[r], K, ib, lb);
            for (int64_t bi = 0; bi < batch; ++bi) {
                float sv = sino[bi * n_ray + r];
                for (int64_t s = 0; s < k; ++s) {
                    double c = (double)(sv * lb[s]);
#pragma omp atomic
                    out[bi * V + ib[s]] += c;
                }
            }
        }
        free(ib);
        free(lb);
    }
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* max over the four body diagonals d of ||M d|| (3d_torch.py:27-35) as torch
 * evaluates it on the CPU in float32: M @ diags with fma-chain dot products,
 * torch.norm(dim=0) with fma-chain squares and sqrtf, then the maximum. */
float orc_voxel_diag_len(const float* M) {
    static const float D[4][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}};
    float best = 0.0f;
    for (int k = 0; k < 4; ++k) {
        float v[3];
        for (int a = 0; a < 3; ++a) {
            float acc = M[a * 3 + 0] * D[k][0];
            acc = fmaf(M[a * 3 + 1], D[k][1], acc);
            acc = fmaf(M[a * 3 + 2], D[k][2], acc);
            v[a] = acc;
        }
        float n = sqrtf(fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0])));
        if (n > best) best = n;
    }
    return best;
}
