/*
 * mpf_hip.h -- the C ABI of the MI355X (gfx950) particle-grid library, libmpfhip.so.
 *
 * Plain pointers, sizes and scalars; no torch or HIP types in the signatures (the stream
 * is an opaque pointer, a hipStream_t, 0 for the null stream).  Every pointer to marker,
 * field or edge data is a DEVICE pointer; geometry arrays of the generic entry points are
 * HOST arrays of ndim entries.  Each call is asynchronous on the given stream and returns
 * MPF_OK or an error code, with the message in mpf_last_error().
 *
 * The entry points replace the reference's kernel catalogue (max-models/mpifield,
 * src/mpifield/kernels/catalog.py:77-114) behind its kernel convention
 * (src/mpifield/weighting.py:22-38, src/mpifield/edges.py:23-31):
 *
 *   scatter(positions, values, field, *geometry[, degree])      adds into field
 *   gather(positions, field, result, *geometry[, degree])        overwrites result
 *   deposit_paths(start, end, charges, edges, *geometry[, degree])  adds into edges
 *
 * with geometry = (*lower, *cell_size, *num_nodes, *code), code 0 bounded (stencils fold
 * onto the wall node), 1 periodic, 2 clamped spline knots (_types.py:27-44).
 *
 * Layouts (weighting.py:31-35): positions (ndim, N) and values/results (ncomp, N) are
 * component-major rows of `pos_stride` / `val_stride` / `result_stride` ELEMENTS apart
 * (rows may have gaps, storage[:, :n]); fields are C-contiguous num_nodes + (ncomp,)
 * float64; edge fields num_nodes (+ degree - 1 on a clamped axis) + (ndim,) float64.
 * Marker arrays are float32 or float64 (`dtype`, MPF_F32 / MPF_F64); cell selection is
 * always done in float64 exactly as the reference does it.
 *
 * Deposit modes:
 *   MPF_MODE_ATOMIC  float64 atomic adds (the plain GPU form of the reference's CUDA
 *                    kernels, e.g. scatter_quadratic_3d_cuda.cu:30-59); order-dependent
 *                    in the last bits.
 *   MPF_MODE_FIXED   each contribution rounded to an int64 fixed-point value of scale
 *                    2^k and summed in integers: bit-reproducible under any marker order.
 *                    k is chosen on the device so that N * max|value| * 2^k < 2^61
 *                    (MPF_SCALE_AUTO) or fixed by the caller.
 *
 * A marker whose values are all zero deposits nothing and its position is not read
 * (weighting.py:34-36); gathers read every position (weighting.py:220-222).
 */
#ifndef MPF_HIP_H
#define MPF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpf_context mpf_context; /* scratch owned by the library (accumulators, bins) */
typedef void *mpf_stream;               /* a hipStream_t */

#define MPF_OK 0
#define MPF_ERR_ARGUMENT 1
#define MPF_ERR_HIP 2
#define MPF_ERR_UNSUPPORTED 3

#define MPF_F32 0
#define MPF_F64 1

#define MPF_MODE_ATOMIC 0
#define MPF_MODE_FIXED 1
#define MPF_SCALE_AUTO (-100000)

#define MPF_KIND_LINEAR 0
#define MPF_KIND_QUADRATIC 1
#define MPF_KIND_SPLINE 2

/* ---- library --------------------------------------------------------------------- */

/* The message of the last failed call on this thread ("" if none). */
const char *mpf_last_error(void);
/* ABI version, bumped on any signature change. */
int mpf_abi_version(void);
/* A context on `device` (its scratch grows on demand); NULL on failure.  A NULL context
 * argument anywhere below means the library's default context of the (current device,
 * stream) pair, made on first use: calls on different streams never share scratch.
 *
 * One stream per context: a context's scratch carries state between the launches of a call
 * (the fixed-point accumulator, the scale slot, bins made inside a call), so calls that
 * share a context must be issued on one stream, or ordered by the caller.  A call that
 * needs more scratch than its context holds synchronizes that context's stream and
 * reallocates; such a call cannot be captured in a HIP graph, so run the work once on the
 * capture stream before capturing it. */
mpf_context *mpf_context_create(int device);
void mpf_context_destroy(mpf_context *ctx);

/* Kernel timing for benchmarks: while enabled, the library records HIP events on the
 * launch stream around its main kernels ("scatter_direct", "gather_direct",
 * "scatter_tiled", "gather_tiled"); mpf_profile_query waits for them and returns the
 * summed milliseconds and the number of timed launches of one name. */
int mpf_profile_enable(int on);
int mpf_profile_reset(void);
int mpf_profile_query(const char *name, double *total_ms, int64_t *launches);

/* ---- generic node kernels -------------------------------------------------------- */

/* kind MPF_KIND_*, degree 0..3 (spline only), ndim 1..3; lower/cell_size/num_nodes/code
 * are host arrays of ndim entries.  Replaces every scatter_* kernel of catalog.py.
 * A fixed-point deposit with the automatic scale of 2^18 to 2^31 - 1 markers, on a shape the
 * tiled kernels take (2D; 3D with whole periodic tiles, 1 or 3 components), is binned by
 * stencil cell inside the call (bins in the context) and deposited tile by tile: the same
 * integers as the direct kernels. */
int mpf_scatter(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, const void *values,
                int64_t val_stride, int dtype, int64_t n, int64_t ncomp, double *field, const double *lower,
                const double *cell_size, const int64_t *num_nodes, const int64_t *code, int mode, int scale_exp,
                mpf_context *ctx, mpf_stream stream);

/* The node deposit of mpf_scatter that allocates nothing.  It has no counterpart in the
 * reference; it is for callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same arguments and results as
 * mpf_scatter (any kind, degree, dimension, boundary code, stride, dtype and mode; always
 * the direct kernels, never the in-call binning), with no context: a fixed-point call keeps
 * its int64 accumulator and scale slot in a static device array of the library, one per GPU
 * and held until the process ends, and takes at most MPF_NOALLOC_MAX_VALUES field values
 * (extent product x ncomp; MPF_ERR_ARGUMENT beyond, before any device work).  Atomic mode
 * uses no scratch.  Every call on one GPU shares that GPU's array, which each call clears
 * before use: fixed-point calls that overlap on one GPU (on two streams at once) race.  The
 * first call on a device looks the array's address up (hipGetSymbolAddress); make it outside
 * graph capture. */
#define MPF_NOALLOC_MAX_VALUES (1 << 23)
int mpf_scatter_noalloc(int kind, int degree, int ndim, const void *positions, int64_t pos_stride,
                        const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp, double *field,
                        const double *lower, const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                        int mode, int scale_exp, mpf_stream stream);

/* Launch variants of the float32 3D tiled deposit and gather (no counterpart in the
 * reference; its kernels have one fixed thread policy, scatter_quadratic_3d_pyccel.py:22-25).
 * A variant is a workgroup tile of 1 or 2 bin tiles along z (8 x 8 x 16 or 8 x 8 x 32 cells:
 * the bins stay the same) and a workgroup of 256 or 512 work-items; the results are the same
 * in every variant.  mpf_tiled_variant sets the variant of the deposit (op 0) or the gather
 * (op 1) for a kind, degree and group of 1 or 3 components (the autotuner's pick) and
 * returns the previous one (-1 on a bad argument; variant < 0 only queries);
 * mpf_tiled_variant_shape describes variant `variant`.  A variant that does not fit a grid
 * (a z axis shorter than its tile) falls back to variant 0. */
int mpf_tiled_variant(int op, int kind, int degree, int ncomp_group, int variant);
int mpf_tiled_variant_shape(int variant, int *tile, int *block);

/* The smallest batch that mpf_scatter and mpf_gather bin inside the call (2^18 markers by
 * default; the per-name catalogue symbols forward to them); returns the previous value, and
 * a negative argument only queries.  A process-wide setting: a huge value keeps every call
 * on the direct, one-marker-per-work-item kernels. */
int64_t mpf_set_binned_min(int64_t markers);

/* Replaces every gather_* kernel of catalog.py; result has the marker dtype.  Batches of
 * 2^18 to 2^31 - 1 markers on the tiled shapes are binned inside the call, gathered tile by
 * tile and written back in the input order. */
int mpf_gather(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, int dtype, int64_t n,
               const double *field, int64_t ncomp, void *result, int64_t result_stride, const double *lower,
               const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
               mpf_stream stream);

/* The derivative-spline gather for differentiation: out[(d * ncomp + c) * out_stride + m]
 * = d/dx_d of the interpolant of component c at marker m (the weights' derivatives along
 * axis d times the others', over the cell size; zero where a bounded axis clamps the
 * position).  No counterpart in the reference, which has no differentiation; the adjoint
 * identity it tests (test_catalog_and_pyccel.py:160-182) is what the backward of the
 * gather and scatter rest on.  Codes 0/1 only. */
int mpf_gather_derivative(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, int dtype,
                          int64_t n, const double *field, int64_t ncomp, void *out, int64_t out_stride,
                          const double *lower, const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                          mpf_context *ctx, mpf_stream stream);

/* ---- the catalogue's node kernels, argument for argument ------------------------- */

#define MPF_SCATTER_IN \
  const void *positions, int64_t pos_stride, const void *values, int64_t val_stride, int dtype, int64_t n, \
      int64_t ncomp, double *field
#define MPF_GATHER_IN \
  const void *positions, int64_t pos_stride, int dtype, int64_t n, const double *field, int64_t ncomp, void *result, \
      int64_t result_stride
#define MPF_GEOM_1D double lower_x, double cell_size_x, int64_t num_x, int64_t periodic_x
#define MPF_GEOM_2D                                                                                               \
  double lower_x, double lower_y, double cell_size_x, double cell_size_y, int64_t num_x, int64_t num_y, \
      int64_t periodic_x, int64_t periodic_y
#define MPF_GEOM_3D                                                                                          \
  double lower_x, double lower_y, double lower_z, double cell_size_x, double cell_size_y, double cell_size_z, \
      int64_t num_x, int64_t num_y, int64_t num_z, int64_t periodic_x, int64_t periodic_y, int64_t periodic_z
#define MPF_TAIL mpf_context *ctx, mpf_stream stream

/* kernels/scatter_linear_{1,2,3}d (weighting.py:259-336: the 1D kernel one component at a
 * time, the 3D ones on rows (nx*ny*nz, ncomp) -- the same memory as num_nodes + (ncomp,)) */
int mpf_scatter_linear_1d(MPF_SCATTER_IN, MPF_GEOM_1D, int mode, MPF_TAIL);
int mpf_scatter_linear_2d(MPF_SCATTER_IN, MPF_GEOM_2D, int mode, MPF_TAIL);
int mpf_scatter_linear_3d(MPF_SCATTER_IN, MPF_GEOM_3D, int mode, MPF_TAIL);
/* kernels/gather_linear_{1,2,3}d */
int mpf_gather_linear_1d(MPF_GATHER_IN, MPF_GEOM_1D, MPF_TAIL);
int mpf_gather_linear_2d(MPF_GATHER_IN, MPF_GEOM_2D, MPF_TAIL);
int mpf_gather_linear_3d(MPF_GATHER_IN, MPF_GEOM_3D, MPF_TAIL);
/* kernels/scatter_quadratic_{1,2,3}d (scatter_quadratic_3d_pyccel.py:130-218) */
int mpf_scatter_quadratic_1d(MPF_SCATTER_IN, MPF_GEOM_1D, int mode, MPF_TAIL);
int mpf_scatter_quadratic_2d(MPF_SCATTER_IN, MPF_GEOM_2D, int mode, MPF_TAIL);
int mpf_scatter_quadratic_3d(MPF_SCATTER_IN, MPF_GEOM_3D, int mode, MPF_TAIL);
/* kernels/gather_quadratic_{1,2,3}d (gather_quadratic_3d_pyccel.py:117-186) */
int mpf_gather_quadratic_1d(MPF_GATHER_IN, MPF_GEOM_1D, MPF_TAIL);
int mpf_gather_quadratic_2d(MPF_GATHER_IN, MPF_GEOM_2D, MPF_TAIL);
int mpf_gather_quadratic_3d(MPF_GATHER_IN, MPF_GEOM_3D, MPF_TAIL);
/* kernels/scatter_spline_{1,2,3}d, degree 0..3 last (scatter_spline_3d_pyccel.py:107-123) */
int mpf_scatter_spline_1d(MPF_SCATTER_IN, MPF_GEOM_1D, int degree, int mode, MPF_TAIL);
int mpf_scatter_spline_2d(MPF_SCATTER_IN, MPF_GEOM_2D, int degree, int mode, MPF_TAIL);
int mpf_scatter_spline_3d(MPF_SCATTER_IN, MPF_GEOM_3D, int degree, int mode, MPF_TAIL);
/* kernels/gather_spline_{1,2,3}d */
int mpf_gather_spline_1d(MPF_GATHER_IN, MPF_GEOM_1D, int degree, MPF_TAIL);
int mpf_gather_spline_2d(MPF_GATHER_IN, MPF_GEOM_2D, int degree, MPF_TAIL);
int mpf_gather_spline_3d(MPF_GATHER_IN, MPF_GEOM_3D, int degree, MPF_TAIL);

/* ---- path deposits (edges.py:187-215) -------------------------------------------- */

#define MPF_PATHS_IN                                                                                           \
  const void *start, int64_t start_stride, const void *end, int64_t end_stride, const void *charges, int dtype, \
      int64_t n, double *edges

/* WHITNEY: kernels/deposit_paths_{1,2,3}d (deposit_paths_2d_pyccel.py:141-183) */
int mpf_deposit_paths_1d(MPF_PATHS_IN, MPF_GEOM_1D, int mode, MPF_TAIL);
int mpf_deposit_paths_2d(MPF_PATHS_IN, MPF_GEOM_2D, int mode, MPF_TAIL);
int mpf_deposit_paths_3d(MPF_PATHS_IN, MPF_GEOM_3D, int mode, MPF_TAIL);
/* QUADRATIC_EDGES / CUBIC_EDGES: kernels/deposit_paths_spline_{1,2,3}d, degree 1..3
 * (deposit_paths_spline_2d_pyccel.py:19-142) */
int mpf_deposit_paths_spline_1d(MPF_PATHS_IN, MPF_GEOM_1D, int degree, int mode, MPF_TAIL);
int mpf_deposit_paths_spline_2d(MPF_PATHS_IN, MPF_GEOM_2D, int degree, int mode, MPF_TAIL);
int mpf_deposit_paths_spline_3d(MPF_PATHS_IN, MPF_GEOM_3D, int degree, int mode, MPF_TAIL);

/* ---- edge gathers (edges.py:169-185), the transpose of the path deposits ---------- */

#define MPF_EDGES_GATHER_IN \
  const void *positions, int64_t pos_stride, int dtype, int64_t n, const double *edges, void *result, \
      int64_t result_stride

/* WHITNEY: kernels/gather_edges_{1,2,3}d (gather_edges_2d_pyccel.py:37-63); result (ndim, n) */
int mpf_gather_edges_1d(MPF_EDGES_GATHER_IN, MPF_GEOM_1D, MPF_TAIL);
int mpf_gather_edges_2d(MPF_EDGES_GATHER_IN, MPF_GEOM_2D, MPF_TAIL);
int mpf_gather_edges_3d(MPF_EDGES_GATHER_IN, MPF_GEOM_3D, MPF_TAIL);
/* QUADRATIC_EDGES / CUBIC_EDGES: kernels/gather_edges_spline_{1,2,3}d, degree 1..3
 * (gather_edges_spline_2d_pyccel.py:52-82) */
int mpf_gather_edges_spline_1d(MPF_EDGES_GATHER_IN, MPF_GEOM_1D, int degree, MPF_TAIL);
int mpf_gather_edges_spline_2d(MPF_EDGES_GATHER_IN, MPF_GEOM_2D, int degree, MPF_TAIL);
int mpf_gather_edges_spline_3d(MPF_EDGES_GATHER_IN, MPF_GEOM_3D, int degree, MPF_TAIL);

/* Generic edge gather: kind 0 Whitney, 1 spline edges of `degree`; result (ndim, n) in
 * the marker dtype with row stride result_stride, overwritten. */
int mpf_gather_edges(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, int dtype, int64_t n,
                     const double *edges, void *result, int64_t result_stride, const double *lower,
                     const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                     mpf_stream stream);

/* Generic form: kind 0 Whitney, 1 spline edges of `degree`. */
int mpf_deposit_paths(int kind, int degree, int ndim, const void *start, int64_t start_stride, const void *end,
                      int64_t end_stride, const void *charges, int dtype, int64_t n, double *edges,
                      const double *lower, const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                      int mode, int scale_exp, mpf_context *ctx, mpf_stream stream);

/* The path deposit of paths the caller keeps binned by their start (2D/3D, bounded or
 * periodic axes): sorted_start (ndim rows of n) and start (the bin offsets) from
 * mpf_bin_markers of the start points -- kind 0 (Whitney) reads bins of the linear cell
 * (kind 0, degree 1), kind 1 bins of the spline nodes of its degree (kind 2) -- with end
 * (ndim rows, row stride end_stride) and charges in the same order.  The fixed-point
 * deposit of mpf_deposit_paths with the automatic scale, its tiled kernel without the sort
 * (the same edges bit for bit).  No counterpart in the reference, whose deposit_paths_block
 * (edges.py:187-215) takes paths in any order: for callers that keep their markers sorted
 * by cell between steps, as the banded CPU buffers assume (scatter_linear_3d_pyccel.py:11-16). */
int mpf_deposit_paths_binned(int kind, int degree, int ndim, const void *sorted_start, const void *end,
                             int64_t end_stride, const void *charges, int dtype, int64_t n, const uint32_t *start,
                             double *edges, const double *lower, const double *cell_size, const int64_t *num_nodes,
                             const int64_t *code, mpf_context *ctx, mpf_stream stream);

/* ---- forms of the spline sequence (splines.py:180-201 faces, cells) ---------------- */

/* kernels/scatter_forms_spline_{1,2,3}d and gather_forms_spline_{1,2,3}d
 * (scatter_forms_spline_3d_pyccel.py:124-210, gather_forms_spline_2d_pyccel.py:93-150):
 * the `form`-form (1 .. ndim) of spline degree 1..3; field num_nodes + (ncomp,) with
 * ncomp = ndim choose form (num + degree - 1 on a clamped axis), values / result
 * (ncomp, n).  Component c is staggered (B_{p-1} shifted half a cell) along its edge
 * (form 1), across its normal (form 2 in 3D) or along every axis (form == ndim), B_p
 * elsewhere; no measure factors, so the two are transposes.  A marker whose values are all
 * zero deposits nothing. */
int mpf_scatter_forms_spline_1d(MPF_SCATTER_IN, MPF_GEOM_1D, int degree, int form, int mode, MPF_TAIL);
int mpf_scatter_forms_spline_2d(MPF_SCATTER_IN, MPF_GEOM_2D, int degree, int form, int mode, MPF_TAIL);
int mpf_scatter_forms_spline_3d(MPF_SCATTER_IN, MPF_GEOM_3D, int degree, int form, int mode, MPF_TAIL);
int mpf_gather_forms_spline_1d(MPF_GATHER_IN, MPF_GEOM_1D, int degree, int form, MPF_TAIL);
int mpf_gather_forms_spline_2d(MPF_GATHER_IN, MPF_GEOM_2D, int degree, int form, MPF_TAIL);
int mpf_gather_forms_spline_3d(MPF_GATHER_IN, MPF_GEOM_3D, int degree, int form, MPF_TAIL);

/* Generic forms (the arrays of mpf_scatter / mpf_gather). */
int mpf_scatter_forms(int degree, int form, int ndim, const void *positions, int64_t pos_stride, const void *values,
                      int64_t val_stride, int dtype, int64_t n, int64_t ncomp, double *field, const double *lower,
                      const double *cell_size, const int64_t *num_nodes, const int64_t *code, int mode,
                      int scale_exp, mpf_context *ctx, mpf_stream stream);
int mpf_gather_forms(int degree, int form, int ndim, const void *positions, int64_t pos_stride, int dtype, int64_t n,
                     const double *field, int64_t ncomp, void *result, int64_t result_stride, const double *lower,
                     const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                     mpf_stream stream);
/* The forms deposit of markers binned by mpf_bin_markers (kind MPF_KIND_SPLINE, the same
 * degree, 2D or 3D, no clamped axis): sorted_positions (ndim, n) and sorted_values
 * (ncomp, n) contiguous rows, start the bin offsets and scalars the fixed-point scale it
 * wrote.  One workgroup per tile of anchor cells, int64 fixed point in LDS: the field is
 * the MODE_FIXED mpf_scatter_forms deposit of the same markers, bit for bit.  The same
 * call as scatter_forms_spline_*d on markers a particle code keeps sorted. */
int mpf_scatter_forms_binned(int degree, int form, int ndim, const void *sorted_positions, const void *sorted_values,
                             int dtype, int64_t n, int64_t ncomp, const void *start, const double *scalars,
                             double *field, const double *lower, const double *cell_size, const int64_t *num_nodes,
                             const int64_t *code, mpf_context *ctx, mpf_stream stream);
/* The forms gather of the same binned markers: result (ncomp, n) in bins order, the values
 * of gather_forms_spline_*d at those markers bit for bit (the tile's field footprint staged
 * in LDS; markers in the dead bin gathered one per work-item). */
int mpf_gather_forms_binned(int degree, int form, int ndim, const void *sorted_positions, int dtype, int64_t n,
                            const void *start, const double *field, int64_t ncomp, void *result, const double *lower,
                            const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                            mpf_stream stream);

/* ---- halos, gradient, divergence -------------------------------------------------- */

/* Halo fill (Field.fill_halos -> mpiarray update_halos, field.py:601-613) and fold
 * (Field.exchange_halos -> accumulate_halos, field.py:615-630) of ONE block holding the
 * whole periodic axis: storage_shape (ndim entries) includes `halo` nodes on every side,
 * data is storage_shape + (ncomp,) float64.  Axis by axis over the full storage, so the
 * corners reach the diagonal in two steps.  Fill: periodic halos take the opposite
 * interior layers, bounded ones are left unchanged.  Fold: periodic halos are added into
 * the opposite interior layers; every halo layer is then cleared. */
int mpf_halo_fill(int ndim, const int64_t *storage_shape, int64_t ncomp, int64_t halo, const int64_t *periodic,
                  double *data, mpf_stream stream);
int mpf_halo_fold(int ndim, const int64_t *storage_shape, int64_t ncomp, int64_t halo, const int64_t *periodic,
                  double *data, mpf_stream stream);

/* EdgeWeighting.divergence_block (edges.py:240-259): out[num_nodes] = G^T edges. */
int mpf_divergence(int ndim, const int64_t *shape, const int64_t *periodic, const double *edges, double *out,
                   mpf_stream stream);
/* EdgeWeighting.gradient_block (edges.py:218-237): out[num_nodes + (ndim,)] = G nodes. */
int mpf_gradient(int ndim, const int64_t *shape, const int64_t *periodic, const double *nodes, double *out,
                 mpf_stream stream);

/* ---- the binned (tiled) path ------------------------------------------------------- */

/* The number of bins of a grid: its stencil cells (product of num_nodes) in 1D and 2D; in
 * 3D the cells padded to whole tiles of 8 x 8 x 16.  The offsets array of mpf_bin_markers
 * holds bins + 2 entries. */
int mpf_bin_sizes(int ndim, const int64_t *num_nodes, int64_t *cells);

/* On-device counting sort of markers by stencil cell (the cell whose nodes a marker's
 * stencil covers, found with the reference's float64 arithmetic): in 3D the cells of each
 * tile of 8 x 8 x 16 cells are consecutive bins (x-major within the tile, tiles x-major),
 * so a tile's markers are one contiguous run; in 1D and 2D, row-major cell order.  A
 * clamped axis (code 2) is binned in 3D for the spline kernels (kind 2) only, by the cell
 * 0 .. num - 2 whose degree + 1 knots' basis functions hold the position.  Markers
 * whose values are all zero go to a trailing bin.  values may be NULL with
 * ncomp 0 (a gather's markers: all live).  Writes, into caller-owned device memory:
 * sorted_positions (ndim, n) and sorted_values (ncomp, n) with contiguous rows, order[n]
 * (int64: sorted index -> input index), start[bins + 2] (uint32 offsets; start[bins] is
 * the first dead marker) and scalars[3] = {2^k, 2^-k, markers per register pass}, the
 * fixed-point scale with n * max|value| * 2^k < 2^61 and max|value| * 2^k < 2^49
 * (scale_exp != MPF_SCALE_AUTO fixes k).  No counterpart in the reference: its banded CPU
 * buffers assume markers sorted by cell (scatter_linear_3d_pyccel.py:11-16) but leave the
 * sort to the caller. */
int mpf_bin_markers(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, const void *values,
                    int64_t val_stride, int dtype, int64_t n, int64_t ncomp, const double *lower,
                    const double *cell_size, const int64_t *num_nodes, const int64_t *code, void *sorted_positions,
                    void *sorted_values, int64_t *order, uint32_t *start, double *scalars, int scale_exp,
                    mpf_context *ctx, mpf_stream stream);

/* Bin again markers kept in the order of earlier bins, after they moved (a pusher updating
 * the sorted arrays in place): positions/values are in the old bins' order, old_order and
 * old_start the old bins' permutation and offsets (same kind, degree and grid).  Writes the
 * outputs of mpf_bin_markers for the same markers -- the same offsets and scale, each bin's
 * markers in some order -- into arrays that must not overlap the inputs, with order[]
 * relative to the markers' first input (old_order composed).  3D: an incremental pass (the
 * markers that stay in their tile of 8 x 8 x 16 cells or reach a neighbouring one) that
 * synchronizes the stream once to read a counter, and sorts from scratch when a marker went
 * further; 1D and 2D, and bins with a clamped axis: mpf_bin_markers.  No counterpart in the reference, whose banded CPU
 * buffers assume markers kept sorted by the caller (scatter_linear_3d_pyccel.py:11-16). */
int mpf_rebin_markers(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, const void *values,
                      int64_t val_stride, int dtype, int64_t n, int64_t ncomp, const int64_t *old_order,
                      const uint32_t *old_start, const double *lower, const double *cell_size,
                      const int64_t *num_nodes, const int64_t *code, void *sorted_positions, void *sorted_values,
                      int64_t *order, uint32_t *start, double *scalars, int scale_exp, mpf_context *ctx,
                      mpf_stream stream);

/* Deposit binned markers (scatter_quadratic_3d etc., the tiled form): one workgroup per
 * tile of cells accumulates int64 fixed point in LDS and adds the tile into field once;
 * bit-reproducible.  3D: every node weighting with 1 to 64 components on bounded axes and
 * the splines' clamped axes of at least a tile (8, 8, 16 cells), periodic axes of at least
 * two tiles and a stencil footprint; 2D, codes 0/1: every node weighting with 1 to 3
 * components; MPF_ERR_UNSUPPORTED otherwise (use mpf_scatter).  item_tables / item_counts: the 3D bins'
 * tables from mpf_bin_items, or NULL (each launch builds them). */
int mpf_scatter_binned(int kind, int degree, int ndim, const void *sorted_positions, const void *sorted_values,
                       int dtype, int64_t n, int64_t ncomp, const uint32_t *start, const uint16_t *item_tables,
                       const uint32_t *item_counts, const double *scalars, double *field, const double *lower,
                       const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                       mpf_stream stream);

/* Gather at binned markers (the tiled form of gather_quadratic_3d etc.); result (ncomp, n)
 * in the sorted order, contiguous rows.  Same cases as mpf_scatter_binned.  item_tables and
 * item_counts are accepted and not read: the gather takes one marker per work-item. */
int mpf_gather_binned(int kind, int degree, int ndim, const void *sorted_positions, int dtype, int64_t n,
                      const uint32_t *start, const uint16_t *item_tables, const uint32_t *item_counts,
                      const double *field, int64_t ncomp, void *result, const double *lower, const double *cell_size,
                      const int64_t *num_nodes, const int64_t *code, mpf_context *ctx, mpf_stream stream);

/* The work lists of 3D bins, made once for bins kept across deposits and gathers:
 * tables[tile * 4096 + i] (uint16: cell | sweep << 12) and counts[tile], tiles = bins / 1024
 * (mpf_bin_sizes).  No counterpart in the reference (whose CPU kernels need no work lists). */
int mpf_bin_items(int ndim, const int64_t *num_nodes, const uint32_t *start, uint16_t *tables, uint32_t *counts,
                  mpf_context *ctx, mpf_stream stream);

/* The gradient of the interpolant at binned markers, contracted with per-marker
 * coefficients: out[d * out_stride + order[m]] = sum_c coef[c * coef_stride + order[m]] *
 * dF_c/dx_d for binned marker m (float64 sums, written in the marker dtype), so that
 * coefficients and result are in the input order.  3D, the cases of mpf_scatter_binned.
 * No counterpart in the reference (which has no differentiation): the tiled form of
 * mpf_gather_derivative followed by the contraction the backward of gather and scatter
 * takes (the adjoint identity of test_catalog_and_pyccel.py:160-182). */
int mpf_gather_derivative_binned(int kind, int degree, int ndim, const void *sorted_positions, int dtype, int64_t n,
                                 const uint32_t *start, const int64_t *order, const double *field, int64_t ncomp,
                                 const void *coef, int64_t coef_stride, void *out, int64_t out_stride,
                                 const double *lower, const double *cell_size, const int64_t *num_nodes,
                                 const int64_t *code, mpf_context *ctx, mpf_stream stream);

/* The fixed-point scale of a deposit of values (ncomp rows of n, row stride val_stride):
 * scalars gets [2^k, 2^-k, markers per register pass] as mpf_bin_markers writes it, so that
 * bins made from positions alone (ncomp 0) can take values later.  No counterpart in the
 * reference; the scale rule is that of the fixed-point deposit (mpf_scatter). */
int mpf_fixed_scale(const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp, int scale_exp,
                    double *scalars, mpf_context *ctx, mpf_stream stream);

/* Rows (rows, n) of binned markers back to the input order: out[c * n + order[m]] =
 * sorted[c * n + m] with order the permutation mpf_bin_markers wrote; both arrays
 * contiguous rows of dtype.  Two rows or more of 2^20 markers or more through a plan made in
 * the stream's default context (mpf_perm_rows), others with scattered writes.  No counterpart in the reference
 * (its markers are never reordered). */
int mpf_unsort_rows(const void *sorted, void *out, const int64_t *order, int64_t rows, int64_t n, int dtype,
                    mpf_stream stream);

/* Marker rows between the input order and the bins' order through a plan of the
 * permutation, in two passes that read and write whole blocks (chunks of 2^15 bins-order
 * indices, buckets of 2^15 input indices, staged through LDS) instead of one pass of
 * scattered accesses.  No counterpart in the reference (its
 * markers are never reordered); for bins kept across steps, whose values arrive and whose
 * results leave in the input order (Field.scatter / Field.gather with bins,
 * Weighting.scatter_block / gather_block with bins).
 *
 * mpf_perm_plan_bytes: the size of the plan of n markers (0 outside 1 .. 2^28).
 * mpf_perm_plan_build: the plan of order (mpf_bin_markers' permutation: bins-order index ->
 *   input index) into device memory of that size (scratch from ctx).
 * mpf_perm_rows: rows of n markers, src rows src_stride elements apart, dst rows dst_stride
 *   apart; direction 0: dst[c * dst_stride + order[m]] = src[c * src_stride + m] (bins order
 *   to input order, as mpf_unsort_rows); direction 1: dst[c * dst_stride + m] =
 *   src[c * src_stride + order[m]] (input order to bins order).  dtype MPF_F32 or MPF_F64;
 *   the staging rows live in ctx. */
int mpf_perm_plan_bytes(int64_t n, int64_t *bytes);
int mpf_perm_plan_build(const int64_t *order, int64_t n, void *plan, mpf_context *ctx, mpf_stream stream);
int mpf_perm_rows(const void *plan, int64_t n, int direction, const void *src, int64_t src_stride, void *dst,
                  int64_t dst_stride, int64_t rows, int dtype, mpf_context *ctx, mpf_stream stream);

/* ---- many patches in one launch ---------------------------------------------------- */

/* One refined-mesh patch: the arguments of one Weighting.scatter_block / gather_block call
 * on BlockGrid.from_geometry(lower, cell_size, num_nodes, periodic) (_types.py:46-96;
 * weighting.py:195-250), with its own weighting (kind, degree) and its slices of the
 * batch's flat marker and field arrays.  Codes 0/1. */
typedef struct mpf_patch {
  double lower[3];
  double cell_size[3];
  int64_t num_nodes[3];
  int64_t code[3];
  int32_t kind;   /* MPF_KIND_* */
  int32_t degree; /* 0..3 for MPF_KIND_SPLINE */
  int64_t first_marker; /* markers [first_marker, first_marker + num_markers) of the batch */
  int64_t num_markers;
  int64_t field_offset; /* the patch field num_nodes + (ncomp,) starts at fields + field_offset */
} mpf_patch;

typedef struct mpf_patch_plan mpf_patch_plan;

/* Upload the patch table and the work list (items of at most `chunk` markers of one
 * patch).  Synchronous; NULL on error.  No counterpart in the reference, where the caller
 * loops over the patches (edges.py:21). */
mpf_patch_plan *mpf_patch_plan_create(int ndim, int64_t n_patches, const mpf_patch *patches, int64_t chunk,
                                      int64_t ncomp);
void mpf_patch_plan_destroy(mpf_patch_plan *plan);
int64_t mpf_patch_plan_items(const mpf_patch_plan *plan);

/* Deposit every patch's markers into its field in one persistent launch of `workgroups`
 * workgroups of `block` (256 or 512) work-items: per item an LDS copy of the patch in
 * int64 fixed point, added into an int64 accumulator, then one finalize pass into the
 * float64 fields.  Bit-reproducible; no host synchronisation. */
int mpf_patches_scatter(const mpf_patch_plan *plan, const void *positions, int64_t pos_stride, const void *values,
                        int64_t val_stride, int dtype, int64_t ncomp, double *fields, int block, int workgroups,
                        mpf_context *ctx, mpf_stream stream);
/* Gather every patch's field at its markers (result (ncomp, N) rows of result_stride). */
int mpf_patches_gather(const mpf_patch_plan *plan, const void *positions, int64_t pos_stride, int dtype,
                       const double *fields, int64_t ncomp, void *result, int64_t result_stride, int workgroups,
                       mpf_context *ctx, mpf_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* MPF_HIP_H */
