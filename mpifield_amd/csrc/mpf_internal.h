// mpf_internal.h -- host-side plumbing shared by the library's translation units:
// the context (scratch owned by the library), error reporting and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/mpf_hip.h"
#include "mpf_device.h"

namespace mpf {

// Scratch that outlives a call: the fixed-point accumulator (kept zero between calls),
// the scale slot, the binning buffers of the tiled path.  Grown on demand, never shrunk.
struct Buffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace mpf

// A context is used by one stream at a time: its scratch carries state from one launch of a
// call to the next (the accumulator, the scale slot, bins made inside a call).  The default
// context of a call with ctx == NULL is the one of its (device, stream) pair.
struct mpf_context {
  int device = 0;
  mpf::Buffer accum;     // int64 fixed-point accumulator, zero between calls
  mpf::Buffer scalars;   // [0] max |v| bits, [1] scale, [2] 1/scale, [3..] spare
  mpf::Buffer sort_keys;
  mpf::Buffer sort_tmp;
  mpf::Buffer offsets;
  mpf::Buffer staging;   // markers in group order while binning
  mpf::Buffer path_in;   // a path deposit's charge and end rows, to be binned with the start
  mpf::Buffer path_sorted;  // the binned paths: start, rows, order, bin offsets, scalars
  mpf::Buffer bins;      // node deposits and gathers binned inside the call: sorted rows, order,
                         // bin offsets, scale, results in bins order
  mpf::Buffer rebin;     // re-binning: markers leaving their tile, per-tile counts and starts
  mpf::Buffer perm_plan;   // the permutation plan of bins made inside a call
  mpf::Buffer perm_build;  // a plan's counts per (chunk, bucket) and their scan down each bucket
  mpf::Buffer perm_stage;  // rows in the staging order between a permutation's two passes
};

namespace mpf {

// The tiles of the binned kernels: 3D anchor cells in tiles of TILE_X x TILE_Y x TILE_Z
// (x-major, z fastest), 2D ones in TILE2 x TILE2 (x-major, y fastest); the bins of a tile
// are consecutive and the tiles follow each other in x-major order.
constexpr int TILE_X = 8, TILE_Y = 8, TILE_Z = 16, TILE_CELLS = TILE_X * TILE_Y * TILE_Z;
constexpr int TILE2 = 32, TILE2_CELLS = TILE2 * TILE2;

void set_error(const char *fmt, ...);
int fail_hip(hipError_t e, const char *what);

// Clear device memory with a kernel (graph-capturable, unlike hipMemsetAsync).
int zero_async(void *p, size_t bytes, hipStream_t stream);

// Grow a buffer to at least `bytes`; `zero` clears new memory.
int ensure(Buffer &b, size_t bytes, bool zero, hipStream_t stream);

// The library's context for (current device, stream), made on first use.
mpf_context *default_context(hipStream_t stream);

// Node deposits and gathers at least this large (mpf_set_binned_min; 2^18 by default) whose
// shape the tiled kernels take are binned inside mpf_scatter / mpf_gather (the bins in the
// context) and run tiled.
constexpr int64_t BINNED_MIN_MARKERS = 1 << 18;
int64_t binned_min();
// Whether mpf_scatter / mpf_gather route a call of this shape through bins and tiles.
bool binned_route(int kind, int degree, int ndim, int64_t ncomp, const Grid &g, int64_t n);
// The node deposit (fixed point, automatic scale) and the node gather through bins made in
// ctx->bins: the same results as mpf_scatter_binned / mpf_gather_binned on the caller's bins.
int scatter_via_bins(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, const void *values,
                     int64_t val_stride, int dtype, int64_t n, int64_t ncomp, double *field, const double *lower,
                     const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                     hipStream_t stream);
int gather_via_bins(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, int dtype, int64_t n,
                    const double *field, int64_t ncomp, void *result, int64_t result_stride, const double *lower,
                    const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                    hipStream_t stream);

// Marker rows between the input order and the bins' order in two passes of whole blocks
// (mpf_permute.hip).  A plan of the permutation `order` (bins-order index -> input index)
// for 1 to 2^28 markers (perm_plan_ok), built into caller memory with the context's
// scratch; direction 0 moves rows from the bins' order to the input order, 1 back.
constexpr int64_t PERM_MIN = 1 << 20;  // the in-call unsort plans from this many markers on
bool perm_plan_ok(int64_t n);
int perm_plan_build(const int64_t *order, int64_t n, void *plan, mpf_context *ctx, hipStream_t stream);
int perm_apply(const void *plan, int64_t n, int direction, const void *src, int64_t src_stride, void *dst,
               int64_t dst_stride, int64_t rows, int dtype, mpf_context *ctx, hipStream_t stream);
// The unsort of rows through a plan made in ctx: out[c * out_stride + order[m]] = sorted[c * n + m].
int unsort_planned(const void *sorted, void *out, int64_t out_stride, const int64_t *order, int64_t rows, int64_t n,
                   int dtype, mpf_context *ctx, hipStream_t stream);

inline int blocks_for(int64_t n, int threads, int64_t cap = 1 << 20) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

// Build the kernel's Grid from the reference's geometry arrays; extent adds degree - 1 on
// clamped axes (edges.py:145-167, scatter_spline_3d_pyccel.py:129).
int make_grid(Grid &g, int ndim, int kind, int degree, const double *lower, const double *cell,
              const int64_t *num, const int64_t *code);

// The fixed-point scale: S = 2^k with k chosen so that n * max|value| * S < 2^61, unless
// the caller fixes k.  Computed on the device into ctx->scalars (no host sync).
int prepare_scale(mpf_context *ctx, const void *values, int64_t vstride, int dtype, int64_t rows, int64_t n,
                  double bound_factor, int scale_exp, hipStream_t stream);
double *scale_slot(mpf_context *ctx);  // device pointer: [0] scale, [1] 1/scale
// The same into caller-given device memory: maxbits (8 bytes) and slot ([scale, 1/scale,
// markers per register pass]); launches only, no allocation.
int scale_into(const void *values, int64_t vstride, int dtype, int64_t rows, int64_t n, double bound_factor,
               int scale_exp, unsigned long long *maxbits, double *slot, hipStream_t stream);
// field[i] += accum[i] * slot[1]; accum[i] = 0 (launches only)
int finalize_into(double *field, unsigned long long *accum, int64_t count, const double *slot, hipStream_t stream);

// field[i] += accum[i] / scale; accum[i] = 0
int finalize_fixed(mpf_context *ctx, double *field, int64_t count, hipStream_t stream);
// the same with the scale read from slot ([scale, 1/scale]) instead of the context's
int finalize_fixed_with(mpf_context *ctx, double *field, int64_t count, const double *slot, hipStream_t stream);

// Kernel timing: when enabled (mpf_profile_enable), the library records a pair of HIP
// events on the launch stream around the launches it names, for mpf_profile_query.
struct ProfileScope {
  const char *name;
  hipStream_t stream;
  int slot;
  ProfileScope(const char *name, hipStream_t stream);
  ~ProfileScope();
};

}  // namespace mpf

#define MPF_CHECK(call)                                           \
  do {                                                            \
    hipError_t _e = (call);                                       \
    if (_e != hipSuccess) return mpf::fail_hip(_e, #call);        \
  } while (0)

#define MPF_LAUNCH_CHECK()                                        \
  do {                                                            \
    hipError_t _e = hipGetLastError();                            \
    if (_e != hipSuccess) return mpf::fail_hip(_e, "kernel launch"); \
  } while (0)
