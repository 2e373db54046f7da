/*
 * ctlab_hip.h -- C ABI of libctlab_hip.so, the MI355X (gfx950) Siddon projector.
 *
 * This library replaces the reference's pybind11 module ct_laboratory._C
 * (src/bindings.cpp:65-104).  Every entry point takes plain device pointers,
 * 64-bit sizes and a hipStream_t (passed as void*; NULL = the default stream),
 * launches asynchronously on that stream and returns a status code.  Outputs are
 * caller-allocated device buffers; the library allocates nothing except where a
 * function says so.  On a nonzero status, ctlab_last_error() returns a message
 * (thread-local, valid until the next call on the same thread).
 *
 * Conventions (as the reference):
 *   volume  [B, n_x, n_y, n_z] float32 C-contiguous, flat = (i*n_y + j)*n_z + k
 *           (src/ct_projector_3d.cu:149-152); 2D image [B, n_row, n_col]
 *           (src/ct_projector_2d.cu:110).
 *   rays    src/dst [n_ray, 3] (or [n_ray, 2]) float32 world coordinates,
 *           P(t) = src + t (dst - src), t in [0, 1].
 *   M_inv   row-major float32 inverse of the voxel->world matrix M (world = M idx + b),
 *           b float32 [3] (or [2]); both device pointers (README.md:111-114).
 *   sinogram [B, n_ray] float32.
 *
 * Semantics flag `mode` (on-the-fly entry points):
 *   CTLAB_MODE_TORCH  segments of ct_projector_{2d,3d}_torch.py: crossing
 *                     t = (p-0.5-s)/dir, length |P(t1)-P(t0)|, no endpoints.
 *   CTLAB_MODE_OTF    segments of the CUDA on-the-fly kernel
 *                     (src/ct_projector_3d.cu:834-1012): t = (p-0.5-s)*(1/dir),
 *                     endpoints 0 and 1 inserted, length (t1-t0)|dst-src|.
 * In both, a segment's voxel is round-half-even of M_inv (midpoint - b).
 *   CTLAB_MODE_MARCH  (2D on-the-fly entry points and ctlab_trace_2d only) the
 *                     incremental march of the reference's 2D on-the-fly kernels
 *                     (src/ct_projector_2d.cu:246-413): entry voxel by floor of
 *                     the entry coordinate, t accumulated per step, length
 *                     (t_hit - t_cur)|dst-src|, every in-grid step visited
 *                     (zero and negative lengths included).
 */
#ifndef CTLAB_HIP_H
#define CTLAB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTLAB_OK 0
#define CTLAB_EINVAL 1   /* bad argument (shape, size, null pointer) */
#define CTLAB_EHIP 2     /* HIP runtime / launch error */
#define CTLAB_ENOTIMPL 3 /* entry point not implemented */

#define CTLAB_MODE_TORCH 0
#define CTLAB_MODE_OTF 1
#define CTLAB_MODE_MARCH 2

#define CTLAB_ZERO_THRESH 0 /* |dir| < 1e-12: axis has no crossings (3D torch, all CUDA kernels) */
#define CTLAB_ZERO_EXACT 1  /* dir == 0: axis has no crossings (ct_projector_2d_torch.py:42) */

typedef void* ctlab_stream_t; /* hipStream_t */

/* Library identification; no counterpart in the reference. */
int ctlab_version(void);
const char* ctlab_last_error(void);

/* ---------------------------------------------------------------------- */
/* 3D on-the-fly (no intersection table)                                   */
/* ---------------------------------------------------------------------- */

/* Replaces _C.forward_project_3d_on_the_fly_cuda (bindings.cpp:98-100;
 * src/ct_projector_3d.cu:1018-1049).  out [batch, n_ray] is overwritten.
 * det_rows/det_cols (0 = none) declare that the rays are ordered
 * [view][det_rows][det_cols] with columns fastest, as the cone-beam builder
 * orders them (conebeam_projector_3d.py:84-96); the kernel then schedules
 * each wavefront over 16 consecutive detector rows of 4 adjacent columns.
 * Results do not depend on it.  With batch >= 2 each ray is traced once for a
 * group of 2 or 4 volumes (same results as one volume at a time). */
int ctlab_forward_project_3d_otf(const float* volume, int64_t batch, int64_t n_x, int64_t n_y, int64_t n_z,
                                 const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                 const float* b, int mode, int64_t det_rows, int64_t det_cols, float* out,
                                 ctlab_stream_t stream);

/* Replaces _C.back_project_3d_on_the_fly_cuda (bindings.cpp:101-103;
 * src/ct_projector_3d.cu:1051-1079).  out [batch, n_x, n_y, n_z]: zeroed first
 * unless accumulate != 0, in which case the projection is added to it. */
int ctlab_back_project_3d_otf(const float* sino, int64_t batch, int64_t n_ray, const float* src,
                              const float* dst, const float* M_inv, const float* b, int64_t n_x, int64_t n_y,
                              int64_t n_z, int mode, int64_t det_rows, int64_t det_cols, int accumulate,
                              float* out, ctlab_stream_t stream);

/* ---------------------------------------------------------------------- */
/* 2D on-the-fly                                                           */
/* ---------------------------------------------------------------------- */

/* Replaces _C.forward_project_2d_on_the_fly_cuda (bindings.cpp:75-77;
 * src/ct_projector_2d.cu:416-440).  zero_rule selects the no-crossing test. */
int ctlab_forward_project_2d_otf(const float* image, int64_t batch, int64_t n_row, int64_t n_col,
                                 const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                 const float* b, int mode, int zero_rule, float* out, ctlab_stream_t stream);

/* Replaces _C.back_project_2d_on_the_fly_cuda (bindings.cpp:78-80;
 * src/ct_projector_2d.cu:442-465).  out [batch, n_row, n_col]: zeroed first
 * unless accumulate != 0.  Computed over 128x64 image tiles summed in LDS in
 * 64-bit fixed point (no global atomic per segment); when the rays are split
 * into several parts per tile (more than 4096 rays and few tiles), the parts
 * combine with one float atomic add per pixel, so the last bits then depend on
 * their order. */
int ctlab_back_project_2d_otf(const float* sino, int64_t batch, int64_t n_ray, const float* src,
                              const float* dst, const float* M_inv, const float* b, int64_t n_row,
                              int64_t n_col, int mode, int zero_rule, int accumulate, float* out,
                              ctlab_stream_t stream);

/* ---------------------------------------------------------------------- */
/* FP64 values over the same traversal (no counterpart in the reference;   */
/* for derivative checks): the segments, voxel indices and float32 lengths */
/* are those of the entry points above, the volume / image, sinogram and   */
/* sums are double.  Forward: one fused multiply-add per segment in        */
/* segment order; back: double atomic adds (out zeroed first unless        */
/* accumulate != 0).  One ray per lane in storage order.                   */
/* ---------------------------------------------------------------------- */
int ctlab_forward_project_3d_otf_f64(const double* volume, int64_t batch, int64_t n_x, int64_t n_y, int64_t n_z,
                                     const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                     const float* b, int mode, double* out, ctlab_stream_t stream);
int ctlab_back_project_3d_otf_f64(const double* sino, int64_t batch, int64_t n_ray, const float* src,
                                  const float* dst, const float* M_inv, const float* b, int64_t n_x, int64_t n_y,
                                  int64_t n_z, int mode, int accumulate, double* out, ctlab_stream_t stream);
int ctlab_forward_project_2d_otf_f64(const double* image, int64_t batch, int64_t n_row, int64_t n_col,
                                     const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                     const float* b, int mode, int zero_rule, double* out, ctlab_stream_t stream);
int ctlab_back_project_2d_otf_f64(const double* sino, int64_t batch, int64_t n_ray, const float* src,
                                  const float* dst, const float* M_inv, const float* b, int64_t n_row,
                                  int64_t n_col, int mode, int zero_rule, int accumulate, double* out,
                                  ctlab_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Per-ray segment lists (parity instrument; no counterpart in the         */
/* reference): for each ray the flat voxel index and length of every        */
/* segment with length > 0, in traversal order.  idx/len are [n_ray, cap];  */
/* counts [n_ray] receives the full count even when it exceeds cap.         */
/* ---------------------------------------------------------------------- */
int ctlab_trace_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* src, const float* dst, int64_t n_ray,
                   const float* M_inv, const float* b, int mode, int zero_rule, int64_t cap, int32_t* idx,
                   float* len, int32_t* counts, ctlab_stream_t stream);
int ctlab_trace_2d(int64_t n_row, int64_t n_col, const float* src, const float* dst, int64_t n_ray,
                   const float* M_inv, const float* b, int mode, int zero_rule, int64_t cap, int32_t* idx,
                   float* len, int32_t* counts, ctlab_stream_t stream);


/* ---------------------------------------------------------------------- */
/* Crossing tables                                                         */
/* ---------------------------------------------------------------------- */

/* Replaces _C.compute_intersections_2d / _3d (bindings.cpp:67-68, 83-84;
 * src/ct_projector_2d.cu:158-177, src/ct_projector_3d.cu:381-401): out [n_ray, K]
 * receives each ray's crossings t in [0,1] in increasing order, then INF.
 * K = (n_row+1)+(n_col+1) (resp. (n_x+1)+(n_y+1)+(n_z+1)) reproduces the
 * reference's full table; a smaller K truncates (rows with more crossings lose
 * their tail).  zero_rule: CTLAB_ZERO_THRESH as the CUDA kernels,
 * CTLAB_ZERO_EXACT as compute_intersections_2d_torch (2d_torch.py:42). */
int ctlab_compute_intersections_2d(int64_t n_row, int64_t n_col, const float* src, const float* dst,
                                   int64_t n_ray, const float* M_inv, const float* b, int zero_rule, int64_t K,
                                   float* out, ctlab_stream_t stream);
int ctlab_compute_intersections_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* src, const float* dst,
                                   int64_t n_ray, const float* M_inv, const float* b, int zero_rule, int64_t K,
                                   float* out, ctlab_stream_t stream);

/* Number of crossings per ray (int32 [n_ray]), for sizing trimmed tables (the
 * reference trims all-INF columns on the host, 3d_module.py:92-98). */
int ctlab_count_crossings_2d(int64_t n_row, int64_t n_col, const float* src, const float* dst, int64_t n_ray,
                             const float* M_inv, const float* b, int zero_rule, int32_t* counts,
                             ctlab_stream_t stream);
int ctlab_count_crossings_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* src, const float* dst,
                             int64_t n_ray, const float* M_inv, const float* b, int zero_rule, int32_t* counts,
                             ctlab_stream_t stream);

/* Replaces _C.forward_project_2d_cuda (bindings.cpp:69-70; src/ct_projector_2d.cu:179-200):
 * tvals [n_ray, K] sorted crossings; pairs with an INF end are skipped. */
int ctlab_forward_project_2d_tvals(const float* image, int64_t batch, int64_t n_row, int64_t n_col,
                                   const float* tvals, int64_t K, const float* src, const float* dst, int64_t n_ray,
                                   const float* M_inv, const float* b, float* out, ctlab_stream_t stream);
/* Replaces _C.back_project_2d_cuda (bindings.cpp:71-72; src/ct_projector_2d.cu:202-221). */
int ctlab_back_project_2d_tvals(const float* sino, int64_t batch, const float* tvals, int64_t K, const float* src,
                                const float* dst, int64_t n_ray, const float* M_inv, const float* b, int64_t n_row,
                                int64_t n_col, int accumulate, float* out, ctlab_stream_t stream);
/* Replaces _C.forward_project_3d_cuda (bindings.cpp:85-86; src/ct_projector_3d.cu:403-433). */
int ctlab_forward_project_3d_tvals(const float* volume, int64_t batch, int64_t n_x, int64_t n_y, int64_t n_z,
                                   const float* tvals, int64_t K, const float* src, const float* dst, int64_t n_ray,
                                   const float* M_inv, const float* b, float* out, ctlab_stream_t stream);
/* Replaces _C.back_project_3d_cuda (bindings.cpp:87-88; src/ct_projector_3d.cu:435-462). */
int ctlab_back_project_3d_tvals(const float* sino, int64_t batch, const float* tvals, int64_t K, const float* src,
                                const float* dst, int64_t n_ray, const float* M_inv, const float* b, int64_t n_x,
                                int64_t n_y, int64_t n_z, int accumulate, float* out, ctlab_stream_t stream);

/* ---------------------------------------------------------------------- */
/* uint16-compressed tables                                                */
/* ---------------------------------------------------------------------- */

#define CTLAB_COMPRESS_CUDA 0  /* compress_tvals_3d_kernel: roundf, INF -> delta 0 (3d.cu:160-212) */
#define CTLAB_COMPRESS_TORCH 1 /* compress_tvals_to_uint16: round-half-even, INF -> 10.0 (3d_torch.py:4-65) */

/* Replaces _C.compress_tvals_3d_cuda (bindings.cpp:90-91; src/ct_projector_3d.cu:464-489):
 * tvals [n_ray, K] -> u16 [n_ray, K] (delta counts, first 0), start, scale [n_ray]. */
int ctlab_compress_tvals_3d(const float* tvals, int64_t n_ray, int64_t K, const float* src, const float* dst,
                            float voxel_diag_len, int semantics, uint16_t* u16, float* start, float* scale,
                            ctlab_stream_t stream);
/* Replaces _C.forward_project_3d_compressed_cuda (bindings.cpp:92-93;
 * src/ct_projector_3d.cu:491-522): t_k = start + (sum of deltas) * scale. */
int ctlab_forward_project_3d_u16(const float* volume, int64_t batch, int64_t n_x, int64_t n_y, int64_t n_z,
                                 const uint16_t* u16, const float* start, const float* scale, int64_t K,
                                 const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                 const float* b, float* out, ctlab_stream_t stream);
/* Replaces _C.back_project_3d_compressed_cuda (bindings.cpp:94-95; src/ct_projector_3d.cu:524-553). */
int ctlab_back_project_3d_u16(const float* sino, int64_t batch, const uint16_t* u16, const float* start,
                              const float* scale, int64_t K, const float* src, const float* dst, int64_t n_ray,
                              const float* M_inv, const float* b, int64_t n_x, int64_t n_y, int64_t n_z,
                              int accumulate, float* out, ctlab_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Table back projection in LDS bricks (same result as                    */
/* ctlab_back_project_3d_tvals / ctlab_back_project_3d_u16 up to the       */
/* grouping of float sums; the reference's consumers add every segment     */
/* with a global atomicAdd, src/ct_projector_3d.cu:215-269, 325-376)       */
/* ---------------------------------------------------------------------- */

/* The volume is cut into bricks of X x Y x Z voxels, brick (i, j, k) holding
 * voxels [32i, 32i+32) x [32j, 32j+32) x [16k, 16k+16), numbered
 * (i * nb_y + j) * nb_z + k with nb_a = ceil(n_a / brick_a). */
#define CTLAB_TABLE_BRICK_X 32
#define CTLAB_TABLE_BRICK_Y 32
#define CTLAB_TABLE_BRICK_Z 16

/* A run: segments seg0 .. seg0+nseg-1 of ray `ray` (segment j runs from table
 * entry j-1 to entry j), every visited one in the same brick; k0 is the uint16
 * walk's running sum at entry seg0-1 (0 for float tables). */
typedef struct ctlab_table_run {
    uint32_t ray, seg0, nseg, k0;
} ctlab_table_run_t;

/* A work unit: runs first .. first+n-1, all of brick `brick`. */
typedef struct ctlab_table_unit {
    uint32_t brick, n;
    int64_t first;
} ctlab_table_unit_t;

/* Number of bricks of an n_x x n_y x n_z grid. */
int64_t ctlab_table_brick_count(int64_t n_x, int64_t n_y, int64_t n_z);

/* Geometry-only run lists of a crossing table: pass exactly one of `tvals`
 * (float [n_ray, K]) and `u16` (uint16 [n_ray, K] with start and scale [n_ray],
 * CUDA decode).  The segments, voxels and skip rules are those of
 * ctlab_back_project_3d_tvals / ctlab_back_project_3d_u16 (ZERO_THRESH grid).
 *   1. ctlab_table_runs_count_3d: counts[brick] = number of runs (the call
 *      clears counts first).
 *   2. The caller forms offsets[brick] = exclusive prefix sum of counts and
 *      n_runs = their total, and allocates runs [n_runs].
 *   3. ctlab_table_runs_fill_3d places each run at offsets[brick] + a cursor
 *      (the call clears cursor [n_brick] first); order within a brick is
 *      unspecified.  A record whose slot would be >= n_runs is dropped. */
int ctlab_table_runs_count_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* tvals, const uint16_t* u16,
                              const float* start, const float* scale, int64_t K, const float* src, const float* dst,
                              int64_t n_ray, const float* M_inv, const float* b, uint32_t* counts,
                              ctlab_stream_t stream);
int ctlab_table_runs_fill_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* tvals, const uint16_t* u16,
                             const float* start, const float* scale, int64_t K, const float* src, const float* dst,
                             int64_t n_ray, const float* M_inv, const float* b, const int64_t* offsets,
                             uint32_t* cursor, ctlab_table_run_t* runs, int64_t n_runs, ctlab_stream_t stream);
/* Back projection [batch, n_x, n_y, n_z] of sino [batch, n_ray] over the runs:
 * one block per (unit, batch item) accumulates in an LDS brick and adds the
 * brick's nonzero voxels to out with global atomics.  The table, rays, map and
 * grid must be those the runs were built from; out is cleared first unless
 * accumulate.  Same extension of the reference as ctlab_back_project_3d_u16 /
 * ctlab_back_project_3d_tvals, whose calls it replaces when the run lists are
 * at hand (bindings.cpp:87-88, 94-95). */
int ctlab_back_project_3d_table_brick(const float* sino, int64_t batch, const float* tvals, const uint16_t* u16,
                                      const float* start, const float* scale, int64_t K, const float* src,
                                      const float* dst, int64_t n_ray, const float* M_inv, const float* b,
                                      int64_t n_x, int64_t n_y, int64_t n_z, const ctlab_table_run_t* runs,
                                      int64_t n_runs, const ctlab_table_unit_t* units, int64_t n_units,
                                      int accumulate, float* out, ctlab_stream_t stream);

/* Forward projection [batch, n_ray] from a crossing table (tvals, or u16 with
 * start and scale, as for the run lists) over each ray's visited range only:
 * segments first[r] .. end[r]-1 (first = end = 0: none), k_first[r] the uint16
 * running sum at entry first[r]-1 (unused for float tables).  The ranges come
 * from the table's run lists (the smallest first segment and the largest end
 * of a ray's runs).  No segment outside its range is visited, so out equals
 * ctlab_forward_project_3d_u16 / ctlab_forward_project_3d_tvals bit for bit
 * (same segments, same order, same fused multiply-adds); whose calls it
 * replaces when the ranges are at hand (bindings.cpp:85-86, 92-93). */
int ctlab_forward_project_3d_table_ranged(const float* volume, int64_t batch, int64_t n_x, int64_t n_y, int64_t n_z,
                                          const float* tvals, const uint16_t* u16, const float* start,
                                          const float* scale, int64_t K, const int32_t* first, const int32_t* end,
                                          const uint32_t* k_first, const float* src, const float* dst,
                                          int64_t n_ray, const float* M_inv, const float* b, float* out,
                                          ctlab_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Atomic-free back projection for flat-panel view stacks (no reference    */
/* counterpart; an alternative schedule of back_project_3d_on_the_fly with  */
/* identical segments)                                                      */
/* ---------------------------------------------------------------------- */

#define CTLAB_FRAME_FLOATS 16

/* Per-view detector frames for rays ordered [n_view][det_rows][det_cols]:
 * frames [n_view, 16] floats (source, pixel (0,0), row and column steps in
 * voxel coordinates, error bound, fit deviations, valid flag).  A view is
 * valid when its rays share one source and its pixels lie on an affine grid
 * within a quarter pixel.  Computed once per geometry. */
int ctlab_cone_frames_3d(const float* src, const float* dst, int64_t n_view, int64_t det_rows, int64_t det_cols,
                         const float* M_inv, const float* b, int64_t n_x, int64_t n_y, int64_t n_z, float* frames,
                         ctlab_stream_t stream);

/* Brick plan of a geometry: for every 32x32x8 brick of the grid and every
 * view, the rectangle of detector pixels that can reach the brick (the
 * brick's shadow) and the running total of rectangle sizes; then the total
 * per brick, each view's midpoint error bound and a copy of the frames.
 * ctlab_brick_plan_ints_3d returns the int32 size of the plan buffer. */
int64_t ctlab_brick_plan_ints_3d(int64_t n_x, int64_t n_y, int64_t n_z, int64_t n_view);
int ctlab_brick_plan_3d(const float* frames, int64_t n_view, int64_t det_rows, int64_t det_cols, int64_t n_x,
                        int64_t n_y, int64_t n_z, int32_t* plan, ctlab_stream_t stream);

/* Same segments as ctlab_back_project_3d_otf (every segment, the same voxel
 * and length, each contribution sino*len), computed brick by brick: each
 * workgroup accumulates one brick in LDS over all views and writes it once, so
 * no global atomics are used.  Sums are 64-bit fixed point scaled per brick
 * (quantum 2^-62 of the brick's bound on a voxel sum), so results do not depend
 * on the order of the rays; bricks whose rays carry INF/NaN use float atomics.
 * Views whose frame is invalid are traced over their whole detector (slower,
 * still exact).  out [batch, n_x, n_y, n_z] is overwritten (or added to when
 * accumulate != 0).  Alternative schedule of the back projection that
 * replaces _C.back_project_3d_on_the_fly_cuda (bindings.cpp:101-103) for
 * rays in the [view][rows][cols] layout.  view_max is caller workspace of
 * batch * n_view floats (each view's largest |sino|, written by the call). */
int ctlab_back_project_3d_brick(const float* sino, int64_t batch, const float* src, const float* dst,
                                int64_t n_view, int64_t det_rows, int64_t det_cols, const int32_t* plan,
                                const float* M_inv, const float* b, int64_t n_x, int64_t n_y, int64_t n_z, int mode,
                                int accumulate, float* out, float* view_max, ctlab_stream_t stream);

/* Filed unit lists of a brick plan (no reference counterpart: they let the
 * brick back projection above, the alternative schedule of
 * _C.back_project_3d_on_the_fly_cuda, bindings.cpp:101-103, skip its per-call
 * filing).  A work unit is one brick and one half of the views (views
 * [0, n_view/2) and [n_view/2, n_view)); its list holds the rays of its
 * shadow rectangles that pass the brick kernel's slab test, in chunks of V
 * consecutive views (V = max(8, ceil(ceil(n_view/2) / 64))), each chunk's
 * rays grouped into 64 eighth-octave classes by estimated crossing count,
 * longest class first in even chunks and last in odd ones.  With
 * S = ctlab_brick_filed_slots_3d(...) / (2 * n_brick) slots per unit, slot
 * (unit * S + c * 64 + j) holds class 63 - j of view chunk c (class j when c
 * is odd) of unit = brick * 2 + half:
 *   1. ctlab_brick_filed_count_3d writes counts[n_slot] (n_slot from
 *      ctlab_brick_filed_slots_3d);
 *   2. the caller forms offsets[n_slot + 1], the exclusive prefix sums of the
 *      counts (int64; the total must stay below 2^32);
 *   3. ctlab_brick_filed_fill_3d writes each slot's ray ids to
 *      ids[offsets[s] .. offsets[s+1]) (order within a slot depends on
 *      timing; n_ids bounds every write).
 * The lists depend only on the rays, layout, grid, map and plan: build them
 * once per plan.  ctlab_back_project_3d_brick_filed then gives the result of
 * ctlab_back_project_3d_brick bit for bit (the same rays per unit, the same
 * per-brick scale), each wave taking the next 64 rays of its unit and
 * skipping rays whose value is zero.  order (or NULL) is the launch order of
 * the units, a permutation of 0 .. n_brick*2 - 1 (of 0 .. n_brick - 1 when
 * accumulate != 0, where a brick's two halves form one unit); longest unit
 * first shortens the launch's tail.  view_max as for the brick kernel. */
int64_t ctlab_brick_filed_slots_3d(int64_t n_x, int64_t n_y, int64_t n_z, int64_t n_view);
int ctlab_brick_filed_count_3d(const float* src, const float* dst, int64_t n_view, int64_t det_rows,
                               int64_t det_cols, const int32_t* plan, const float* M_inv, const float* b, int64_t n_x,
                               int64_t n_y, int64_t n_z, uint32_t* counts, ctlab_stream_t stream);
int ctlab_brick_filed_fill_3d(const float* src, const float* dst, int64_t n_view, int64_t det_rows,
                              int64_t det_cols, const int32_t* plan, const float* M_inv, const float* b, int64_t n_x,
                              int64_t n_y, int64_t n_z, const int64_t* offsets, uint32_t* ids, int64_t n_ids,
                              ctlab_stream_t stream);
int ctlab_back_project_3d_brick_filed(const float* sino, int64_t batch, const float* src, const float* dst,
                                      int64_t n_view, int64_t det_rows, int64_t det_cols, const int32_t* plan,
                                      const int64_t* offsets, const uint32_t* ids, const int32_t* order,
                                      const float* M_inv, const float* b, int64_t n_x, int64_t n_y, int64_t n_z,
                                      int mode, int accumulate, float* out, float* view_max, ctlab_stream_t stream);

/* Brick back projection of rays in any order (no reference counterpart: an
 * alternative schedule of _C.back_project_3d_on_the_fly_cuda, bindings.cpp:101-103,
 * for rays that are not a stack of flat-panel views).  A binning pass lists, per
 * 32x32x8 brick, the rays that can have a segment in it (the rays passing within
 * E + 0.05 voxel of the brick's cells, E the ray's midpoint error bound):
 *   1. ctlab_brick_bins_count_3d writes counts[n_brick] (rays per brick) and
 *      emax[1] (the largest E over rays with segments);
 *   2. the caller forms offsets[n_brick + 1], the exclusive prefix sums of the
 *      counts, and a copy of offsets[0..n_brick) as cursor;
 *   3. ctlab_brick_bins_fill_3d writes the ray ids of brick k to
 *      ids[offsets[k] .. offsets[k+1]) (in an order that depends on timing;
 *      n_ids, the length of ids, bounds every write; fill ids with 0xffffffff
 *      first: ids of n_ray or more are no ray);
 *   4. ctlab_brick_bins_file_3d refiles each brick's list in place order into
 *      filed[offsets[k] .. offsets[k+1]): two units per brick (rays of even and
 *      of odd id), each in chunks of consecutive ids, each chunk grouped by 64
 *      eighth-octave classes of estimated crossing count, longest first in
 *      even chunks and last in odd ones; slot_offsets
 *      [n_slot + 1] (int64, n_slot from ctlab_brick_bins_slots_3d) holds the
 *      start of every slot, slot (brick * 2 + parity) * S + chunk * 64 + j with
 *      S = n_slot / (2 * n_brick).
 * The lists depend only on the rays, grid, map and rule: build them once per ray
 * set.  ctlab_back_project_3d_brick_bins then computes the same segments as
 * ctlab_back_project_3d_otf brick by brick in LDS (64-bit fixed point, no global
 * atomics, results independent of the order within the lists), each wave
 * taking the next 64 rays of its unit; order (or NULL) is the launch order of
 * the units, a permutation of 0 .. n_brick*2 - 1 (of 0 .. n_brick - 1 when
 * accumulate != 0); vmax is caller workspace of batch floats.  n_ray must be
 * below 2^32.  ctlab_brick_count_3d returns n_brick. */
int64_t ctlab_brick_count_3d(int64_t n_x, int64_t n_y, int64_t n_z);
int64_t ctlab_brick_bins_slots_3d(int64_t n_x, int64_t n_y, int64_t n_z, int64_t n_ray);
int ctlab_brick_bins_count_3d(const float* src, const float* dst, int64_t n_ray, const float* M_inv, const float* b,
                              int64_t n_x, int64_t n_y, int64_t n_z, int mode, uint32_t* counts, float* emax,
                              ctlab_stream_t stream);
int ctlab_brick_bins_fill_3d(const float* src, const float* dst, int64_t n_ray, const float* M_inv, const float* b,
                             int64_t n_x, int64_t n_y, int64_t n_z, int mode, uint32_t* cursor, uint32_t* ids,
                             int64_t n_ids, ctlab_stream_t stream);
int ctlab_brick_bins_file_3d(const float* src, const float* dst, int64_t n_ray, const float* M_inv, const float* b,
                             int64_t n_x, int64_t n_y, int64_t n_z, const uint32_t* offsets, const uint32_t* ids,
                             const float* emax, int64_t* slot_offsets, uint32_t* filed, ctlab_stream_t stream);
int ctlab_back_project_3d_brick_bins(const float* sino, int64_t batch, const float* src, const float* dst,
                                     int64_t n_ray, const int64_t* slot_offsets, const uint32_t* filed,
                                     const float* emax, const int32_t* order, const float* M_inv, const float* b,
                                     int64_t n_x, int64_t n_y, int64_t n_z, int mode, int accumulate, float* out,
                                     float* vmax, ctlab_stream_t stream);

/* Largest batch * n_view of ctlab_back_project_3d_brick_noalloc (256 KiB of
 * per-view maxima, held by each device the library runs on until the process
 * ends). */
#define CTLAB_BRICK_SCRATCH_FLOATS 65536

/* ctlab_back_project_3d_brick without the view_max workspace: the per-view
 * maxima go to a static device array of CTLAB_BRICK_SCRATCH_FLOATS floats, so
 * the call allocates nothing (no hipMalloc, hipFree or allocating library
 * routine).  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Same arguments and results as
 * ctlab_back_project_3d_brick otherwise; batch * n_view above
 * CTLAB_BRICK_SCRATCH_FLOATS is rejected with CTLAB_EINVAL.  Every call on one
 * device shares that device's array (each call writes all the entries it
 * reads), so calls that overlap on one device (different streams, or
 * concurrent host threads) race; serialise them. */
int ctlab_back_project_3d_brick_noalloc(const float* sino, int64_t batch, const float* src, const float* dst,
                                        int64_t n_view, int64_t det_rows, int64_t det_cols, const int32_t* plan,
                                        const float* M_inv, const float* b, int64_t n_x, int64_t n_y, int64_t n_z,
                                        int mode, int accumulate, float* out, ctlab_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Many 2D scans in one persistent launch (no reference counterpart: the   */
/* reference projects scans one launch each, ct_projector_2d_cuda.py:23-46) */
/* ---------------------------------------------------------------------- */

/* One scan.  Pointers are device addresses stored as integers.  For the
 * forward projection `image` is read and `sino` written; for the back
 * projection `sino` is read and `image` written.  item_begin = index of the
 * scan's first work item: scans are listed in item order and scan k owns items
 * [item_begin_k, item_begin_k + ceil(n_ray_k / rays_per_item)). */
typedef struct ctlab_scan2d {
    uint64_t image; /* float [n_row, n_col] */
    uint64_t sino;  /* float [n_ray] */
    uint64_t src;   /* float [n_ray, 2] */
    uint64_t dst;   /* float [n_ray, 2] */
    float minv[4];  /* row-major inverse of M */
    float b[2];
    int32_t n_row, n_col;
    int64_t n_ray;
    int32_t mode;      /* CTLAB_MODE_TORCH or CTLAB_MODE_OTF (any other value takes the torch rule) */
    int32_t zero_rule; /* CTLAB_ZERO_THRESH or CTLAB_ZERO_EXACT */
    int64_t item_begin;
} ctlab_scan2d;

/* Forward projections of all scans: `scans` is a device array of n_scan
 * records, total_items the sum of the scans' item counts, n_groups the number
 * of persistent workgroups, counter 8 bytes of device memory (cleared by the
 * call).  Same segments and results as ctlab_forward_project_2d_otf per scan. */
int ctlab_forward_project_2d_batch(const ctlab_scan2d* scans, int64_t n_scan, int64_t total_items,
                                   int64_t rays_per_item, int64_t n_groups, void* counter, ctlab_stream_t stream);

/* Back projections of all scans into their images (zeroed first unless
 * accumulate != 0). */
int ctlab_back_project_2d_batch(const ctlab_scan2d* scans, int64_t n_scan, int64_t total_items,
                                int64_t rays_per_item, int64_t n_groups, void* counter, int accumulate,
                                ctlab_stream_t stream);

/* Back projections of all scans without global atomics per segment (no
 * counterpart in the reference; same results as ctlab_back_project_2d_batch
 * within float summation order).  Work units are (tile_rows x 64 image tile,
 * tile_rows = 128 or 64, range of rays_per_part consecutive rays): in this call
 * item_begin is the scan's first UNIT, and scan k owns ceil(n_row/tile_rows) *
 * ceil(n_col/64) * ceil(n_ray/rays_per_part) units
 * (ctlab_back_project_2d_tile_units); total_units is their sum.  A unit sums its rays' contributions to its
 * tile in LDS (64-bit fixed point, order independent) and writes the tile
 * once: with plain stores when the scan has one part, else with one float
 * atomic add per pixel.  Images are zeroed first unless accumulate != 0.
 * Units past a scan's last one (item_begin of the next scan, or total_units,
 * set too high) are skipped. */
int ctlab_back_project_2d_batch_tiled(const ctlab_scan2d* scans, int64_t n_scan, int64_t total_units,
                                      int64_t rays_per_part, int64_t tile_rows, int64_t n_groups, void* counter,
                                      int accumulate, ctlab_stream_t stream);

/* Number of work units of one scan in ctlab_back_project_2d_batch_tiled
 * (-1 for invalid sizes or tile_rows other than 64 and 128). */
int64_t ctlab_back_project_2d_tile_units(int64_t n_row, int64_t n_col, int64_t n_ray, int64_t rays_per_part,
                                         int64_t tile_rows);

/* ---------------------------------------------------------------------- */
/* MAP iteration with a Gaussian likelihood and the quadratic smoothness   */
/* prior.  The reference runs it as autograd + torch.optim.SGD            */
/* (bayesian_estimation/map_estimator.py:113-128, random_variable/        */
/* gaussian.py:54-63, random_variable/image_energies.py:18-29); these      */
/* kernels are the parts around the projectors, fused, with no host sync.  */
/* ---------------------------------------------------------------------- */

/* r = (ax - y) * inv_var; energy[0] += sum 0.5 (ax - y)^2 inv_var (if energy). */
int ctlab_map_residual(const float* ax, const float* y, int64_t n, float inv_var, float* r, double* energy,
                       ctlab_stream_t stream);

/* x_out = x - lr * d2 * (bp + beta * L x) on the grid [n0, n1, n2] (n2 fastest;
 * 2D images use n0 = 1), L = sum over axes of diff^T diff with free ends
 * (the gradient of (beta/2) sum ||torch.diff(x, dim)||^2); d2 = NULL means 1.
 * energy[0] += (beta/2) sum ||diff||^2 of x (if energy).  x_out != x. */
int ctlab_map_prior_update(const float* x, const float* bp, const float* d2, float beta, float lr, int64_t n0,
                           int64_t n1, int64_t n2, float* x_out, double* energy, ctlab_stream_t stream);

/* Separable quadratic surrogate weights: d2 = 1 / (ata1 * inv_var + 2 beta deg),
 * ata1 = A^T A 1, deg = number of grid neighbours of the voxel. */
int ctlab_map_sqs_weights(const float* ata1, float inv_var, float beta, int64_t n0, int64_t n1, int64_t n2,
                          float* d2, ctlab_stream_t stream);

/* p[0..n) = 0 with a kernel (n <= 256), usable inside a captured graph. */
int ctlab_zero_f64(double* p, int64_t n, ctlab_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CTLAB_HIP_H */
