/*
 * pygrog_hip.h -- C ABI of libpygrog_hip.so, the MI355X (gfx950) GROG
 * gridding library.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a nonzero code on error;
 *     pg_last_error() then returns a message for the calling thread.
 *   - Pointers are device pointers unless stated otherwise.  Sizes are
 *     element counts.  Complex arrays are interleaved (re, im) pairs:
 *     "c64" = float2, "c128" = double2.
 *   - `stream` is a hipStream_t (NULL = legacy default stream).  All work is
 *     enqueued asynchronously on it; the entry points that end with a small
 *     host read (pg_plan_compile, pg_group_by_entry) synchronise that stream
 *     and say so, no other entry point waits on the device, and no entry
 *     point allocates device memory: outputs and workspaces are
 *     caller-provided.
 *
 * The first block mirrors the seven ops of the reference's pybind module
 * `_pygrog_torch` (csrc/torch/module.cpp:54-87), one entry per dtype.  The
 * remaining blocks are the device-resident plan compiler, the fused gridding
 * kernels and the helpers they need; their Python callers live in
 * pygrog_amd/_lib.py and pygrog_amd/calib/_grog.py.
 */
#ifndef PYGROG_HIP_H
#define PYGROG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *pg_stream_t; /* hipStream_t */

/* ------------------------------------------------------------------------
 * Library status
 * ---------------------------------------------------------------------- */
/* Message of the last failed call on this thread ("" if none). */
const char *pg_last_error(void);
/* ABI version of this header (bumped on any signature change). */
int pg_abi_version(void);

/* ------------------------------------------------------------------------
 * Reference native ops (csrc/torch/module.cpp:54-87)
 * ---------------------------------------------------------------------- */

/* out[n,b,:] = kernel[indexes[n]] @ data[n,b,:]
 * data, out: (N, B, C); indexes: (N,) int64; kernel: (K, C, C) row-major.
 * An index outside [0, K) yields a zero output row.
 * Replaces grog_interpolate (csrc/torch/grog_interp.cpp:112-149;
 * CUDA grog_interp_cuda.cu:28-106). */
int pg_grog_interpolate_c64(const void *data, const int64_t *indexes,
                            const void *kernel, void *out, int64_t N,
                            int64_t B, int64_t C, int64_t K,
                            pg_stream_t stream);
int pg_grog_interpolate_c128(const void *data, const int64_t *indexes,
                             const void *kernel, void *out, int64_t N,
                             int64_t B, int64_t C, int64_t K,
                             pg_stream_t stream);

/* grid[indices[i]] += weights[i] * data[i], in place; indices outside
 * [0, grid_size) are skipped.  Order-fixed: the first element of each run of
 * equal consecutive indices adds the run's terms one by one, in input order,
 * to the cell's current value (each product and sum rounded separately) and
 * commits the result with a compare-and-swap.  With sorted indices -- the
 * reference CPU op's contract -- that is exactly the reference's arithmetic
 * (sparse_ops_cpu_impl.inl:77-120), bit for bit on every call.  Unsorted
 * input stays correct: a run that loses the swap to another run of the same
 * cell adds its partial sum atomically (order then not fixed).
 * Replaces scatter_add (csrc/torch/sparse_ops.cpp:233-267). */
int pg_scatter_add_c64(void *grid, int64_t grid_size, const void *data,
                       const int64_t *indices, const float *weights,
                       int64_t n, pg_stream_t stream);
int pg_scatter_add_c128(void *grid, int64_t grid_size, const void *data,
                        const int64_t *indices, const double *weights,
                        int64_t n, pg_stream_t stream);
/* pg_scatter_add_c64 on `rows` grids at once (rows <= 65535): grid (rows,
 * grid_size) and data (rows, n) share indices and weights; row r is
 * pg_scatter_add_c64(grid[r], data[r]), the same bits.  Serves the per-coil
 * loops of the shot-wise gridding (reference _grog.py:697-731) and of
 * interpolate(grid=True) (_grog.py:763-821) in complex128. */
int pg_scatter_add_rows_c64(void *grid, int64_t grid_size, const void *data,
                            const int64_t *indices, const float *weights,
                            int64_t n, int64_t rows, pg_stream_t stream);
int pg_scatter_add_rows_c128(void *grid, int64_t grid_size, const void *data,
                             const int64_t *indices, const double *weights,
                             int64_t n, int64_t rows, pg_stream_t stream);

/* Same contract as pg_scatter_add_*; `bin_starts` (n_bin_starts int64, the
 * searchsorted bin edges of the reference) is validated but not needed by
 * the run fold.
 * Replaces scatter_add_binned (csrc/torch/sparse_ops.cpp:299-340). */
int pg_scatter_add_binned_c64(void *grid, int64_t grid_size, const void *data,
                              const int64_t *indices, const float *weights,
                              int64_t n, const int64_t *bin_starts,
                              int64_t n_bin_starts, int64_t bin_size,
                              pg_stream_t stream);
int pg_scatter_add_binned_c128(void *grid, int64_t grid_size,
                               const void *data, const int64_t *indices,
                               const double *weights, int64_t n,
                               const int64_t *bin_starts, int64_t n_bin_starts,
                               int64_t bin_size, pg_stream_t stream);

/* out[i] = weights[i] * grid[indices[i]] (0 where the index is out of range).
 * Replaces gather (csrc/torch/sparse_ops.cpp:269-297). */
int pg_gather_c64(const void *grid, int64_t grid_size, const int64_t *indices,
                  const float *weights, int64_t n, void *out,
                  pg_stream_t stream);
int pg_gather_c128(const void *grid, int64_t grid_size,
                   const int64_t *indices, const double *weights, int64_t n,
                   void *out, pg_stream_t stream);

/* psf[indices[i]] += w_sq[i], in place (real PSF).  The three PSF builders
 * fold like pg_scatter_add_c64: for sorted indices each cell (and each PSF
 * element) equals the reference's in-order CPU sum bit for bit
 * (toep_psf.cpp:58-237); unsorted input stays correct, order not fixed.
 * Replaces psf_scatter_scalar (csrc/torch/toep_psf.cpp:242-274). */
int pg_psf_scatter_scalar_f32(float *psf, int64_t grid_size,
                              const int64_t *indices, const float *w_sq,
                              int64_t n, pg_stream_t stream);
int pg_psf_scatter_scalar_f64(double *psf, int64_t grid_size,
                              const int64_t *indices, const double *w_sq,
                              int64_t n, pg_stream_t stream);

/* psf[indices[i], l, l'] += w_sq[i] * conj(b[i,l]) * b[i,l'];
 * psf: (grid_size, M, M) complex; basis_per_sample: (n, M) complex.
 * Replaces psf_scatter_outer (csrc/torch/toep_psf.cpp:276-320). */
int pg_psf_scatter_outer_c64(void *psf, int64_t grid_size, int64_t M,
                             const int64_t *indices, const float *w_sq,
                             const void *basis_per_sample, int64_t n,
                             pg_stream_t stream);
int pg_psf_scatter_outer_c128(void *psf, int64_t grid_size, int64_t M,
                              const int64_t *indices, const double *w_sq,
                              const void *basis_per_sample, int64_t n,
                              pg_stream_t stream);

/* psf[indices[i], k, k'] += w_sq[i] * conj(B[k,t_i]) * B[k',t_i];
 * psf: (grid_size, K, K) complex; basis: (K, T) complex; time_index: (n,).
 * Samples with t_i outside [0, T) are skipped.
 * Replaces psf_scatter_outer_basis (csrc/torch/toep_psf.cpp:322-367). */
int pg_psf_scatter_outer_basis_c64(void *psf, int64_t grid_size, int64_t K,
                                   const int64_t *indices, const float *w_sq,
                                   const void *basis, int64_t T,
                                   const int64_t *time_index, int64_t n,
                                   pg_stream_t stream);
int pg_psf_scatter_outer_basis_c128(void *psf, int64_t grid_size, int64_t K,
                                    const int64_t *indices,
                                    const double *w_sq, const void *basis,
                                    int64_t T, const int64_t *time_index,
                                    int64_t n, pg_stream_t stream);

/* ------------------------------------------------------------------------
 * Plan compiler (device-resident GrogPlan).  Replaces the host-side
 * _create_plan_single / _attach_fft_plan / calc_interp_table quantisation
 * (src/pygrog/calib/_grog.py:1081-1234, 918-998, 382-422).
 * ---------------------------------------------------------------------- */

/* Geometry shared by the plan kernels (host struct, passed by value). */
typedef struct pg_plan_geom {
  int32_t ndim;           /* 2 or 3 */
  int32_t kw;             /* kernel offsets per source */
  int64_t grid_shape[3];  /* oversampled grid, row-major (axis 0 slowest) */
  int64_t strides[3];     /* flat-index strides of grid_shape */
  float half_amp[3];      /* 0.5 * (2 * (s - 1 - s // 2)), f32 */
  float origins[3];       /* -(s // 2) */
  float steps[3];         /* (s - 1) / (gs - 1), f32 */
  int32_t offsets[128][3]; /* kernel offsets, kw rows (kw <= 128) */
} pg_plan_geom;

/* cmax[s, d] = max_n |coords[s, n, d]|;  coords (S, npts, ndim) f32. */
int pg_plan_absmax(const float *coords, int64_t S, int64_t npts, int32_t ndim,
                   float *cmax, pg_stream_t stream);

/* Per (s, n, k) candidate replica (natural order, nper = npts * kw):
 *   target_idx (S, nper) int64   flat grid index or -1 (invalid);
 *   distances  (S, nper, ndim) f32  target - source in image-grid units;
 *   gauss      (S, nper) f32     exp(-0.5 |d/steps|^2), 0 when invalid;
 *   sort_key   (S*nper) uint64   s*(G+1) + (valid ? cell + 1 : 0);
 *   sort_val   (S*nper) uint32   s*nper + n*kw + k;
 *   n_invalid  (S) int64         per-element invalid counts (must be zeroed
 *                                 by the caller; accumulated atomically).
 * Float work follows the reference's f32 operation order without
 * contraction, so target_idx and distances are bit-exact. */
int pg_plan_targets(const float *coords, const float *cmax, int64_t S,
                    int64_t npts, pg_plan_geom geom, int64_t *target_idx,
                    float *distances, float *gauss, uint64_t *sort_key,
                    uint32_t *sort_val, int64_t *n_invalid,
                    pg_stream_t stream);

/* Stable LSD radix sort of (uint64 key, uint32 value) pairs on bits
 * [0, end_bit).  Call with workspace == NULL to query *ws_bytes. */
int pg_sort_pairs_u64(void *workspace, size_t *ws_bytes,
                      const uint64_t *keys_in, uint64_t *keys_out,
                      const uint32_t *vals_in, uint32_t *vals_out, int64_t n,
                      int32_t end_bit, pg_stream_t stream);
/* Stable radix sort of (uint32 key, uint32 value) pairs on bits [0, end_bit). */
int pg_sort_pairs_u32(void *workspace, size_t *ws_bytes,
                      const uint32_t *keys_in, uint32_t *keys_out,
                      const uint32_t *vals_in, uint32_t *vals_out, int64_t n,
                      int32_t end_bit, pg_stream_t stream);
/* out[i] = sum_{j<i} in[j] (uint32); out has n + 1 entries (out[n] = total).
 * Call with workspace == NULL to query *ws_bytes. */
int pg_exclusive_scan_u32(void *workspace, size_t *ws_bytes,
                          const uint32_t *in, uint32_t *out, int64_t n,
                          pg_stream_t stream);

/* From the sorted (key, value) arrays of pg_plan_targets:
 *   indices (S, nper) int64 sorted grid index (invalid -> 0);
 *   sort_perm, inv_perm (S, nper) int64;
 *   gauss_sorted (S*nper) f32 = gauss[value];
 *   cell_ptr (S, G+1) uint32 CSR of the valid replicas of each element
 *   (replica r of element s = sorted position n_invalid[s] + r). */
int pg_plan_post(const uint64_t *keys_sorted, const uint32_t *vals_sorted,
                 const float *gauss, const int64_t *n_invalid, int64_t S,
                 int64_t nper, int64_t G, int64_t *indices, int64_t *sort_perm,
                 int64_t *inv_perm, float *gauss_sorted, uint32_t *cell_ptr,
                 pg_stream_t stream);

/* EGROG weights (reference _grog.py:1182-1212, 958-975) per cell:
 *   gsum = sum (f64, sorted order) of gauss; w = gauss / f32(max(gsum,1e-12));
 *   weights (S, nper) f32 natural order (0 for invalid);
 *   sqrt_weights (S, nper) f32 sorted order (0 for invalid);
 *   rep_src (S, Rmax) uint32, rep_s (S, Rmax) f32 per valid replica;
 *   density (S, G) f32 = sum of sqrt_w^2 in sorted order;
 *   unit_count (S*G) uint32 = ceil(replicas / chunk). */
int pg_plan_weights(const uint32_t *vals_sorted, const float *gauss_sorted,
                    const int64_t *n_invalid, const uint32_t *cell_ptr,
                    int64_t S, int64_t nper, int64_t G, int32_t kw,
                    int32_t chunk, float *weights, float *sqrt_weights,
                    uint32_t *rep_src, float *rep_s, float *density,
                    uint32_t *unit_count, int64_t rep_stride,
                    pg_stream_t stream);

/* Unit arrays from the exclusive scan of unit_count (cell_unit, S*G+1):
 *   unit_cell (U) uint32 cell of each unit (element-local);
 *   unit_rep (U + S) uint32 first replica of each unit, element-local,
 *   with a terminating entry per element: unit_rep[cell_unit[(s+1)G] + s]. */
int pg_plan_units(const uint32_t *cell_ptr, const uint32_t *cell_unit,
                  int64_t S, int64_t G, int32_t chunk, uint32_t *unit_cell,
                  uint32_t *unit_rep, pg_stream_t stream);

/* Normal-equation products of the Cartesian GRAPPA unit-operator fit
 * (reference calib/_grappa.py:16-59 and 211-287: ridge least squares of
 * each axis' target samples on its source samples), in FP64: calib is
 * (C, *dims) complex128 coil-major, dims (host, ndim = 2 or 3 entries, each
 * >= 2).  out (complex128, 2 ndim C^2 + C entries): for axis d,
 * out[(2d) C^2 + i C + j] = sum_p conj(x_i[p]) x_j[p] and
 * out[(2d+1) C^2 + i C + j] = sum_p conj(x_i[p]) x_j[p + e_d] over the
 * positions p below the last slice along d; out[2 ndim C^2 + i] =
 * sum_p |x_i[p]|^2 (real) over all positions.  The unit operator of axis d
 * for regularisation lamda (with the calibration scaled to unit norm, as
 * the reference does) is then ((G_d + lamda |x|^2 I)^-1 X_d)^T. */
int pg_grappa_gram_f64(const void *calib, int64_t C, int32_t ndim,
                       const int64_t *dims, void *out, pg_stream_t stream);

/* Operator table from the per-axis fractional powers (reference
 * prepare_grappa_table, _grappa.py:131-181): powers (ndim, nsteps, C, C)
 * complex128 = Gx, Gy(, Gz) powers for the nsteps exponents;
 *   2D: T[i nsteps + j] = Gx[j] Gy[i]
 *   3D: T[i nsteps^2 + j nsteps + k] = (F[k] Gy[j]) L[i], (F, L) = (Gx, Gz)
 *       (reference pairing) or (Gz, Gx) when `corrected`,
 * in FP64 (ascending fma chains), written to table_c64 (K, C, C) complex64
 * rounded to nearest and, when table_c128 is not NULL, to table_c128
 * (K, C, C) complex128.  C <= 32. */
int pg_table_compose_f64(const void *powers, int32_t ndim, int64_t nsteps, int64_t C,
                         int32_t corrected, void *table_c64, void *table_c128,
                         pg_stream_t stream);

/* Hot cells and the tile kernels' row numbering from the reduction units.
 * cell_unit (S*G + 1) is the exclusive scan of units per cell over all
 * elements.  A cell with nu units is hot when nu >= min_units and special
 * (no tile-kernel rows) when hot or nu > 1.  pg_plan_hot_counts writes
 * (S*G + 1 entries each, the last 0): a = replicas of ordinary cells (0
 * for special ones), hf = 1 for hot cells, hn = units of hot cells.  With
 * A, HF, HN their exclusive scans (pg_exclusive_scan_u32), pg_plan_hot_rows
 * writes cell_vrow (S, G + 1) = A less the element's base; per element s
 * (hb = HF[s*G], ub = HN[s*G], H_s = HF[(s+1)*G] - hb): hot_ptr[hb + s + h]
 * for h = 0..H_s = first hot unit of the element's h-th hot cell relative
 * to ub (h = H_s: the element's hot-unit count), hot_units[HN[cell] + k] =
 * global unit id of the k-th unit of each hot cell, and unit_hot[u] = the
 * element-local hot index of unit u's cell, -1 when not hot.  hot_ptr
 * holds HF[S*G] + S entries, hot_units HN[S*G], unit_hot cell_unit[S*G].
 * No reference counterpart: the pre-pass work lists of the fused gridding. */
int pg_plan_hot_counts(const uint32_t *cell_ptr, const uint32_t *cell_unit,
                       int64_t S, int64_t G, int64_t min_units, uint32_t *a,
                       uint32_t *hf, uint32_t *hn, pg_stream_t stream);
int pg_plan_hot_rows(const uint32_t *cell_unit, int64_t S, int64_t G,
                     int64_t min_units, const uint32_t *A, const uint32_t *HF,
                     const uint32_t *HN, uint32_t *cell_vrow, int32_t *unit_hot,
                     uint32_t *hot_units, uint32_t *hot_ptr, pg_stream_t stream);

/* The whole plan compiler in one call (the steps above, in order: absmax,
 * targets, stable key sort, post, weights, units, hot counts, scans, hot
 * rows), for callers that want no per-step orchestration (the serving
 * loop's preparation threads).  pg_plan_compile_layout gives the byte
 * offsets of the PG_PLAN_FIELDS arrays inside one output workspace (in
 * this order: target_idx i64 (S, nper), weights f32 (S, nper), indices
 * i64, sqrt_weights f32, sort_perm i64, inv_perm i64, cell_ptr u32
 * (S, G + 1), rep_src u32 (S, max(nper, 1)), rep_s f32 (S, max(nper, 1)),
 * density f32 (S, G), cell_unit u32 (S G + 1), unit_cell u32 (Umax),
 * unit_rep u32 (Umax + S), unit_hot i32 (Umax), hot_units u32 (Umax),
 * hot_ptr u32 (S G + S), cell_vrow u32 (S, G + 1), n_invalid i64 (S),
 * summary i64 (3 S + 3); Umax = max(S nper, 1) bounds the units), its
 * size and the size of the scratch workspace (free once the stream has run
 * the call).  pg_plan_compile enqueues everything on `stream`, writes the
 * candidates' shifts to `distances` (S, nper, ndim) f32 (a buffer of its
 * own, so it can be released alone), then synchronises the stream once and
 * fills the host array summary[3 S + 3] = [n_invalid[s] for s < S, U (the
 * unit count), hot-cell bases HF[s G] for s <= S, hot-unit bases HN[s G]
 * for s <= S] (the scans of pg_plan_hot_rows).  The arrays are those the
 * separate entry points produce, bit for bit.
 * No reference counterpart beyond the steps it runs. */
#define PG_PLAN_FIELDS 19
int pg_plan_compile_layout(int64_t S, int64_t npts, int32_t ndim, int32_t kw,
                           int64_t G, int64_t *offsets, int64_t *out_bytes,
                           int64_t *tmp_bytes);
int pg_plan_compile(const float *coords, int64_t S, int64_t npts,
                    pg_plan_geom geom, int64_t G, int32_t chunk,
                    int64_t min_units, void *out_ws, int64_t out_ws_bytes,
                    void *tmp_ws, int64_t tmp_ws_bytes, float *distances,
                    int64_t *summary, pg_stream_t stream);
/* pg_plan_compile without the synchronisation: the summary is copied into
 * summary_pinned (page-locked host memory, e.g. from hipHostMalloc) by a copy
 * enqueued on `stream`, and is valid once the stream has reached that point
 * (the caller records an event or synchronises; tmp_ws must stay allocated
 * until then).  For callers that overlap host work with the plan kernels
 * (the serving loop fits the job's operator table meanwhile).  No reference
 * counterpart beyond the steps it runs. */
int pg_plan_compile_async(const float *coords, int64_t S, int64_t npts,
                          pg_plan_geom geom, int64_t G, int32_t chunk,
                          int64_t min_units, void *out_ws, int64_t out_ws_bytes,
                          void *tmp_ws, int64_t tmp_ws_bytes, float *distances,
                          int64_t *summary_pinned, pg_stream_t stream);

/* Replicas grouped by table entry for pg_apply_grouped_bf16* (no reference
 * counterpart: the grouped path's preparation).  For the R replicas with
 * entries rep_ent[0..R) < K: ent_rep = the replica ids in a stable sort by
 * entry, ent_ptr (K + 1) each entry's range of positions, items the work list
 * of (entry, first offset) pairs of chunks of `chunk` replicas in entry
 * order, ent_src / ent_s each position's source and weight; *n_items (host)
 * = the work list's length, written after the call has synchronised
 * `stream`.  pg_group_by_entry_layout gives the work list's capacity and the
 * temporary workspace's size (1 <= R < 2^31). */
int pg_group_by_entry_layout(int64_t R, int64_t K, int64_t chunk, int64_t *items_max,
                             int64_t *tmp_bytes);
int pg_group_by_entry(const uint32_t *rep_ent, const uint32_t *rep_src,
                      const float *rep_s, int64_t R, int64_t K, int64_t chunk,
                      uint32_t *ent_ptr, uint32_t *ent_rep, uint32_t *ent_src,
                      float *ent_s, uint32_t *items, int64_t items_max,
                      void *tmp, int64_t tmp_bytes, int64_t *n_items,
                      pg_stream_t stream);
/* pg_group_by_entry without the synchronisation: the work list's slots from
 * its length up to items_max hold entry 0xFFFFFFFF (which pg_apply_grouped_bf16*
 * skips), so the apply can be launched with n_items = items_max before the
 * length is known on the host; the length itself is copied into
 * *n_items_pinned when that is not NULL (page-locked host memory that stays
 * allocated until `stream` has passed this point; zeroed by the call).  For callers that enqueue a job's
 * preparation and go on (the serving loop).  No reference counterpart. */
int pg_group_by_entry_async(const uint32_t *rep_ent, const uint32_t *rep_src,
                            const float *rep_s, int64_t R, int64_t K,
                            int64_t chunk, uint32_t *ent_ptr, uint32_t *ent_rep,
                            uint32_t *ent_src, float *ent_s, uint32_t *items,
                            int64_t items_max, void *tmp, int64_t tmp_bytes,
                            int64_t *n_items_pinned, pg_stream_t stream);

/* Quantised table entry of every candidate replica (reference
 * calc_interp_table, _grog.py:406-422, generalised from 10**precision to an
 * integer quantisation q):  pfac = q, stepsize = f32(step), nsteps = q + 1,
 *   idx_d = rint((radius + rint(d_d * pfac) / pfac) / stepsize),
 *   entry = clamp(sum_d rint(idx_d * nsteps^d), 0, n_entries - 1).
 * interp_idx (S, nper) int64 natural order; rep_ent (S, Rmax) uint32 for
 * valid replicas in cell-sorted order (via inv_perm).
 * axis_order: 0 = reference pairing (entry axis d <- distances[..., d]). */
int pg_plan_quantize(const float *distances, const int64_t *inv_perm,
                     const int64_t *n_invalid, const int64_t *target_idx,
                     int64_t S, int64_t nper, int32_t ndim, float pfac,
                     float stepsize, float radius, int64_t nsteps,
                     int64_t n_entries, int64_t *interp_idx, uint32_t *rep_ent,
                     int64_t rep_stride, pg_stream_t stream);

/* ------------------------------------------------------------------------
 * Fused gridding (reference interpolate(grid=True) -> _grid_kspace,
 * src/pygrog/calib/_grog.py:736-902, on top of grog_interp.cpp:39-73 and
 * sparse_ops_cpu_impl.inl:38-120).
 * ---------------------------------------------------------------------- */

/* One frame = one (source block, plan, table, output grid) job. */
typedef struct pg_frame {
  const void *y;            /* sources: (n_src, C) sample-major c64 for
                               pg_grid_fp32, (C, n_src) coil-major for
                               pg_grid_scatter */
  void *grid;               /* (C, n_cell) c64, fully written */
  const void *table;        /* packed table, pg_table_pack_c64 layout */
  const uint32_t *rep_src;  /* (R) source of each replica, cell-sorted */
  const uint32_t *rep_ent;  /* (R) table entry of each replica */
  const float *rep_s;       /* (R) sqrt weight of each replica */
  const uint32_t *unit_cell;/* (U) */
  const uint32_t *unit_rep; /* (U + 1) */
  const uint32_t *cell_unit;/* (n_cell + 1) first unit of each cell */
  const void *vbuf;         /* (R, C) c64 applied replicas (grouped modes) */
  const int32_t *unit_hot;  /* (U) hot-cell index of a unit, -1 if none */
  const uint32_t *hot_units;/* (n_hot_units) units of the hot cells, in order */
  const uint32_t *hot_ptr;  /* (n_hot + 1) hot cell h owns hot_units[hot_ptr[h]..] */
  void *p_unit;             /* (n_hot_units, C) c64 scratch: hot unit partials */
  const uint32_t *cell_row; /* (n_cell + 1) first replica of each cell (CSR) */
  const uint32_t *cell_vrow;/* (n_cell + 1) replicas of ordinary cells (not
                               hot, one unit) before each cell: the tile
                               kernels' row numbering */
  int64_t n_src;
  int64_t n_cell;
  int64_t n_hot;
  int64_t n_hot_units;
} pg_frame;

/* Pack a (K, C, C) c64 row-major table into the layout read by the fused
 * kernels: packed[e][p][co] = (K[co][2p], K[co][2p+1]) as float4,
 * p < G/2 where G is the coil group of C (the smallest of 4, 8, 16, 32, 64
 * that holds C), zero padded past C.  packed holds
 * pg_table_packed_elems(K, C) = K * (G/2) * C float4. */
int pg_table_pack_c64(const void *table, int64_t K, int64_t C, void *packed,
                      pg_stream_t stream);
int64_t pg_table_packed_elems(int64_t K, int64_t C);

/* Plain FP32 fused gridding: per replica v = (K[e] y[:, src]) * s * s,
 * reduced per cell in plan order (chunked sequential folds), written as a
 * dense coil-major grid (zeros in empty cells).  frames: device array of F
 * pg_frame; tile_ptr: device array (F + 1) of cumulative tile counts with
 * tile_cells cells per tile.  Deterministic: the result does not depend on
 * tile size, block count or the other frames in the launch.
 * Cells with many replicas ("hot", e.g. the radial DC) are reduced by two
 * pre-pass kernels -- unit partials in parallel into p_unit, then an
 * in-order fold per hot cell written straight into the grid -- on a helper
 * stream forked from and joined back into `stream` (graph-capturable), so
 * they overlap the tile kernel, which leaves hot cells alone; a hot cell
 * costs one workgroup no more than any other and the result is the same
 * bits as without the pre-pass.  max_hot_units / max_hot: largest
 * n_hot_units / n_hot over the frames (0 = no pre-pass). */
int pg_grid_fp32(const pg_frame *frames, const uint32_t *tile_ptr, int64_t F,
                 int64_t n_tiles, int64_t C, int64_t tile_cells,
                 int64_t n_blocks, int64_t max_hot_units, int64_t max_hot,
                 pg_stream_t stream);

/* Second pass of the grouped modes: same reduction as pg_grid_fp32 but v is
 * read from frame.vbuf[r] (cell-sorted rows) instead of being computed. */
int pg_grid_reduce_v(const pg_frame *frames, const uint32_t *tile_ptr,
                     int64_t F, int64_t n_tiles, int64_t C,
                     int64_t tile_cells, int64_t n_blocks,
                     int64_t max_hot_units, int64_t max_hot,
                     pg_stream_t stream);
/* pg_grid_reduce_v with frame.vbuf holding block-floating-point rows (the
 * bfp15 layout of pg_bfp_row_bytes(C) bytes per replica, as written by
 * pg_apply_grouped_bf16_bfp): each value is decoded, then folded in the same
 * canonical order.  tile_cells must be the wave tile (256); other tile sizes
 * are an error.  No reference counterpart. */
int pg_grid_reduce_v_bfp(const pg_frame *frames, const uint32_t *tile_ptr,
                         int64_t F, int64_t n_tiles, int64_t C,
                         int64_t tile_cells, int64_t n_blocks,
                         int64_t max_hot_units, int64_t max_hot,
                         pg_stream_t stream);

/* Third mode, used by SparseFFT (reference _scatter_ifft_crop,
 * src/pygrog/operator/_sparse_fft.py:396-435, 1154-1165): identity operator,
 * v = y[:, rep_src[r]] * rep_s[r] with rep_src = sort_perm and rep_s =
 * sqrt_weights, reduced per cell exactly like pg_grid_fp32. */
int pg_grid_scatter(const pg_frame *frames, const uint32_t *tile_ptr,
                    int64_t F, int64_t n_tiles, int64_t C, int64_t tile_cells,
                    int64_t n_blocks, int64_t max_hot_units, int64_t max_hot,
                    pg_stream_t stream);

/* Default cells per tile of the gridding kernels for a coil count (0 if
 * C > 64). */
int pg_grid_tile_cells(int64_t C);

/* The tile sizes compiled for a coil count, largest first, written to
 * opts[0 .. min(n, max_opts)); returns n (0 if C is out of [1, 64]).  A
 * launch's tile_cells must be one of them and its tile_ptr built with it. */
int pg_grid_tile_options(int64_t C, int32_t *opts, int32_t max_opts);

/* out[b, m] = weights[inv_perm[m]] * grid[b, indices[inv_perm[m]]]:
 * SparseFFT gather fused with the unsort (reference _fft_pad_gather,
 * src/pygrog/operator/_sparse_fft.py:1167-1177).  grid (B, G), out (B, n). */
int pg_gather_perm_c64(const void *grid, int64_t G, const int64_t *indices,
                       const float *weights, const int64_t *inv_perm,
                       int64_t n, int64_t B, void *out, pg_stream_t stream);

/* The same gather for a whole batch of stacked plans in one launch (the
 * reference's _fft_pad_gather_batch / stacked _forward_packed,
 * src/pygrog/operator/_sparse_fft.py:626-680, 1050-1149): grid (B, S, C, G),
 * indices / weights / inv_perm (S, n) (inv_perm element-local), out
 * (B, S, C, n); out[b, s, c] is pg_gather_perm_c64 of grid[b, s] with
 * element s's plan, bit for bit. */
int pg_gather_perm_stack_c64(const void *grid, int64_t G, const int64_t *indices,
                             const float *weights, const int64_t *inv_perm,
                             int64_t n, int64_t C, int64_t S, int64_t B,
                             void *out, pg_stream_t stream);

/* unit_count[s*G + c] = ceil((cell_ptr[s][c+1] - cell_ptr[s][c]) / chunk). */
int pg_unit_count(const uint32_t *cell_ptr, int64_t S, int64_t G,
                  int32_t chunk, uint32_t *unit_count, pg_stream_t stream);

/* y (C, N) c64 -> y_t (N, C) c64 sample-major. */
int pg_transpose_cn_c64(const void *y, int64_t C, int64_t N, void *y_t,
                        pg_stream_t stream);

/* Frames of uneven sample counts, packed back to back: frame f is a (C, N_f)
 * coil-major block at sample offset frame_off[f] (element offset
 * C * frame_off[f]); its (N_f, C) sample-major rows are written at the same
 * offset of y_t.  frame_off (F + 1 entries, device) is non-decreasing and
 * max_n >= every N_f.  No reference counterpart: it feeds the batched
 * multi-frame gridding launch. */
int pg_transpose_frames_c64(const void *y, const int64_t *frame_off, int64_t F,
                            int64_t max_n, int64_t C, void *y_t,
                            pg_stream_t stream);

/* 24-bit sample-major layout, the BF16 path's copy of the k-space: a
 * sample's row holds its 2C values (re, im per coil), 3 bytes each -- the
 * top three bytes of the fp32 value rounded to nearest even at bit 8 (16
 * significant bits; NaN stays NaN, overflow rounds to infinity), packed
 * little-endian -- and is padded to pg_s24_row_bytes(C) = 16 * ceil(6C / 16)
 * bytes (-1 if C < 0).  The split-BF16 products use 16 bits of an input
 * (hi = bf16(x), lo = bf16(x - hi)), so the copy holds what they read in 3/4
 * of the bytes.  No reference counterpart: an internal format of the BF16
 * gridding path. */
int64_t pg_s24_row_bytes(int64_t C);
/* pg_transpose_cn_c64 into the 24-bit layout: y_s (16-byte aligned) holds
 * N rows of pg_s24_row_bytes(C) bytes. */
int pg_transpose_cn_s24(const void *y, int64_t C, int64_t N, void *y_s,
                        pg_stream_t stream);
/* pg_transpose_frames_c64 into the 24-bit layout: frame f's rows start at
 * byte pg_s24_row_bytes(C) * frame_off[f] of y_s (16-byte aligned). */
int pg_transpose_frames_s24(const void *y, const int64_t *frame_off, int64_t F,
                            int64_t max_n, int64_t C, void *y_s,
                            pg_stream_t stream);

/* Block-floating-point sample-major layout ("bfp15"), the compact copy of
 * the k-space for the split-BF16 products: a sample's 2C values (re, im per
 * coil) in groups of 8 (4 coils; the last group zero-padded), 16 bytes a
 * group, pg_bfp_row_bytes(C) = 16 * ceil(C / 4) bytes a row (-1 if C < 0):
 * a 32-coil row is one 128-byte line.  A group is 8 little-endian 16-bit
 * fields; bits 1..15 of field i are a two's-complement mantissa m_i in
 * [-16384, 16383], bit 0 of field i is bit i of the group's exponent E, and
 * value_i = m_i * 2^(E - 140).  The encoder takes M = max |x_i| over the
 * group, s = max(floor(log2 M) - 13, -140) (+1 when M * 2^-s rounds past
 * 16383; s = -140 for M = 0), m_i = x_i * 2^-s rounded to nearest even: every
 * value is within M * 2^-14 of its input.  A group holding a NaN or an
 * infinity decodes to -infinity.  No reference counterpart: an internal
 * format of the BF16 gridding path (lossy; the bf16x3 products read every
 * stored value exactly). */
int64_t pg_bfp_row_bytes(int64_t C);
/* pg_transpose_cn_c64 into the bfp15 layout: y_b (16-byte aligned) holds N
 * rows of pg_bfp_row_bytes(C) bytes. */
int pg_transpose_cn_bfp(const void *y, int64_t C, int64_t N, void *y_b,
                        pg_stream_t stream);
/* pg_transpose_frames_c64 into the bfp15 layout: frame f's rows start at
 * byte pg_bfp_row_bytes(C) * frame_off[f] of y_b (16-byte aligned). */
int pg_transpose_frames_bfp(const void *y, const int64_t *frame_off, int64_t F,
                            int64_t max_n, int64_t C, void *y_b,
                            pg_stream_t stream);

/* Pack a (K, C, C) c64 table for the BF16 MFMA grouped GEMM: per entry and
 * part, (re, im) bf16 pairs of K[co][ci] padded to 8*ceil(2C/16) rows by
 * 16*ceil(2C/32) columns; part 0 = hi = bf16(x), part 1 = lo =
 * bf16(x - hi).  The kernel expands each pair into the real 2C x 2C form in
 * registers.  n_parts = 1 or 2; packed holds pg_table_bf16_elems bf16. */
int pg_table_pack_bf16(const void *table, int64_t K, int64_t C, int32_t n_parts,
                       void *packed, pg_stream_t stream);

/* Grouped complex GEMM on BF16 MFMA with FP32 accumulation.  Replicas are
 * grouped by table entry: ent_ptr (K+1) and ent_rep (R) = the cell-sorted
 * replica index at each position of the entry order (both from
 * pg_sort_pairs_u32 on the replicas' entries); ent_src / ent_s (R) = that
 * replica's source sample and sqrt weight, in the same entry order.  items
 * (n_items pairs of uint32: entry, first offset inside its group) are the
 * 128-replica work chunks, one wavefront each; for the replica r at each
 * position, vbuf[r] = (K[e] y_t[src]) * s * s.  n_terms = 1 (hi*hi) or 3
 * (hi*hi + hi*lo + lo*hi split-BF16 products). */
int pg_apply_grouped_bf16(const void *y_t, const void *table_bf16,
                          const uint32_t *ent_ptr, const uint32_t *ent_rep,
                          const uint32_t *ent_src, const float *ent_s,
                          const uint32_t *items, int64_t n_items,
                          int64_t K, int64_t C, int32_t n_terms, void *vbuf,
                          pg_stream_t stream);
/* pg_apply_grouped_bf16 reading the 24-bit sample rows of
 * pg_transpose_cn_s24 / pg_transpose_frames_s24 (y_s 16-byte aligned); the
 * products and the V rows are as there.  No reference counterpart. */
int pg_apply_grouped_bf16_s24(const void *y_s, const void *table_bf16,
                              const uint32_t *ent_ptr, const uint32_t *ent_rep,
                              const uint32_t *ent_src, const float *ent_s,
                              const uint32_t *items, int64_t n_items,
                              int64_t K, int64_t C, int32_t n_terms, void *vbuf,
                              pg_stream_t stream);
/* pg_apply_grouped_bf16 reading the bfp15 sample rows of
 * pg_transpose_cn_bfp / pg_transpose_frames_bfp (y_b 16-byte aligned) and
 * writing V in the same bfp15 row layout (vbuf: replica r's row of
 * pg_bfp_row_bytes(C) bytes at byte r * pg_bfp_row_bytes(C), 16-byte
 * aligned): each 4-coil group of (K[e] y_b[src]) * s * s is encoded as the
 * transposes encode k-space.  The products are as there; reduce with
 * pg_grid_reduce_v_bfp.  No reference counterpart. */
int pg_apply_grouped_bf16_bfp(const void *y_b, const void *table_bf16,
                              const uint32_t *ent_ptr, const uint32_t *ent_rep,
                              const uint32_t *ent_src, const float *ent_s,
                              const uint32_t *items, int64_t n_items,
                              int64_t K, int64_t C, int32_t n_terms, void *vbuf,
                              pg_stream_t stream);
/* Element count (bf16) of the pg_table_pack_bf16 layout, -1 if C > 64. */
int64_t pg_table_bf16_elems(int64_t K, int64_t C, int32_t n_parts);

/* SENSE expand + zero pad + centring sign, one pass (reference
 * SparseFFT._fft_pad_gather / _adjoint_single image side,
 * src/pygrog/operator/_sparse_fft.py:703-789, 1167-1177, and
 * _base/_fftc.py:114-139): out (B, C, *grid) c64,
 * out[b,c,g] = s(g) * smaps[c, g-off] * x[b, g-off] inside the centred image
 * window, 0 outside; smaps NULL: x is (B, C, *image) and out = s(g) * x.
 * s(g) = (-1)^(sum of g's indices) if modulate else 1; off = (grid -
 * image) // 2.  With modulate, fftn(out) times (-1)^(k - grid/2) per axis
 * is the centred FFT of the padded image (even grid sizes).  image_shape,
 * grid_shape: host arrays of ndim (1..3) entries. */
int pg_sense_expand_c64(const void *x, const void *smaps, int64_t B, int64_t C,
                        int32_t ndim, const int64_t *image_shape,
                        const int64_t *grid_shape, int32_t modulate, void *out,
                        pg_stream_t stream);

/* Crop + centring sign + conj(smaps) coil sum, one pass (reference
 * SparseFFT._scatter_ifft_crop / _forward_single combine,
 * src/pygrog/operator/_sparse_fft.py:574-621, 1154-1165): full (B, C, *grid)
 * c64 -> out (B, *image) = sign * sum_c conj(smaps[c]) * s(p+off) *
 * full[b, c, p+off] (coils summed in order); smaps NULL: out (B, C, *image)
 * = sign * s(p+off) * full[b, c, p+off]. */
int pg_sense_combine_c64(const void *full, const void *smaps, int64_t B,
                         int64_t C, int32_t ndim, const int64_t *image_shape,
                         const int64_t *grid_shape, int32_t modulate,
                         float sign, void *out, pg_stream_t stream);

/* out[0] + i out[1] = sum_i conj(a_i) * b_i for complex64 a, b (n values),
 * accumulated in double (the FP64 CG / LSMR scalars), deterministic.
 * workspace: device buffer of pg_vdot_workspace_bytes() bytes; out: device
 * array of 2 doubles. */
int pg_vdot_c64_f64(const void *a, const void *b, int64_t n, void *workspace,
                    double *out, pg_stream_t stream);
int64_t pg_vdot_workspace_bytes(void);

/* One iteration of the CG solver's vector and scalar work (cg with FP64
 * scalars and no preconditioner, src/pygrog/_solve/_solvers.py:49-139),
 * fused: sc (FP64, device) holds [0] rz, [1] alpha, [2] beta, [3] r0,
 * [4] rel, [5] rr.  pg_cg_step_xr: pAp = Re<p, Ap>, alpha = rz / (pAp +
 * 1e-30), x_out = x_in + f32(alpha) p, r -= f32(alpha) Ap, rr = Re<r, r>,
 * rel = sqrt(rr) / r0, beta = rr / (rz + 1e-30), rz = rr (workspace of
 * pg_vdot_workspace_bytes()).  pg_cg_step_p: p = r + f32(beta) p.  The same
 * roundings and sum orders as the separate operations (pg_vdot_c64_f64 for
 * the inner products), so the iterates equal theirs bit for bit. */
int pg_cg_step_xr(const void *p, const void *ap, const void *x_in, void *x_out,
                  void *r, int64_t n, double *sc, void *workspace,
                  pg_stream_t stream);
int pg_cg_step_p(const void *r, void *p, int64_t n, const double *sc,
                 pg_stream_t stream);

/* pg_vdot_c64_f64 without a workspace and without allocating: the same
 * kernels and the same bits, with the block partials in a 16 KiB array the
 * library holds in device memory (one per GPU, kept until the process ends;
 * its address is looked up once per device, with that device current).  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight solver or serving loops).  n is limited to
 * PG_VDOT_NOALLOC_MAX_N (larger n is rejected with an error); the scratch
 * need does not grow with n (at most 1024 partials).  Every call on one GPU
 * shares that GPU's array and leaves its last contents there: calls that
 * overlap on one GPU (different streams, or host threads) race -- serialise
 * them, e.g. on one stream.  The pointers must belong to the current
 * device. */
#define PG_VDOT_NOALLOC_MAX_N ((int64_t)1 << 40)
int pg_vdot_c64_f64_noalloc(const void *a, const void *b, int64_t n,
                            double *out, pg_stream_t stream);

/* interpolate(ret_image=False) fused: out[b, co, n*kw + k] =
 *   K[interp_idx[n, k]] @ y[b, :, n]  for every candidate replica
 * (including invalid ones, as the reference does).  y (B, C, N) c64,
 * interp_idx (N*kw) int64, table (K, C, C) c64 row-major, out (B, C, N*kw).
 * Replaces _sparse_full_single + _interpolate (_grog.py:668-695, 1283-1295). */
int pg_sparse_apply_c64(const void *y, const int64_t *interp_idx,
                        const void *table, void *out, int64_t B, int64_t C,
                        int64_t N, int64_t kw, int64_t K, pg_stream_t stream);

/* pg_sparse_apply_c64 for a batch of a stacked plan in one launch (the
 * reference's per-element loop in _sparse_full, _grog.py:563-666): y
 * (B, S, C, N), interp_idx (S, N*kw), out (B, S, C, N*kw); block (b, s) is
 * pg_sparse_apply_c64 of y[b, s] with element s's entries, bit for bit. */
int pg_sparse_apply_stack_c64(const void *y, const int64_t *interp_idx,
                              const void *table, void *out, int64_t B, int64_t S,
                              int64_t C, int64_t N, int64_t kw, int64_t K,
                              pg_stream_t stream);

/* cell_ptr (G + 1) uint32 from sorted int64 indices (n of them); indices
 * outside [0, G) are not allowed (the caller checks sortedness/range). */
int pg_csr_from_sorted(const int64_t *indices, int64_t n, int64_t G,
                       uint32_t *cell_ptr, pg_stream_t stream);
/* Same for sorted uint32 keys (table-entry groups of the grouped modes). */
int pg_csr_from_sorted_u32(const uint32_t *keys, int64_t n, int64_t G,
                           uint32_t *cell_ptr, pg_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PYGROG_HIP_H */
