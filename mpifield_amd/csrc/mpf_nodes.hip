// mpf_nodes.hip -- the node deposit and gather, one work-item per marker (the general
// path: every kind, degree 0..3, 1-3 dimensions, boundary codes 0/1/2, any number of
// components, strided marker rows, float32 or float64 markers).
//
// Deposit: the reference's kernels (scatter_quadratic_3d_pyccel.py:88-127,
// scatter_spline_3d_pyccel.py:62-104, scatter_linear_{1,2,3}d_pyccel.py) run one marker per
// work-item; each of its (p+1)^ndim x ncomp contributions goes to the field with a
// float64 atomic (MODE_ATOMIC) or, rounded to int64 fixed point, to the context's
// accumulator that a finalize pass adds into the field (MODE_FIXED).
// Gather: gather_*_pyccel.py, one marker per work-item, the weight products and the sum in
// the reference's order.
#include <map>
#include <mutex>

#include "mpf_internal.h"

namespace mpf {

template <class T, int KIND, int NDIM, int MODE>
__global__ void __launch_bounds__(256) scatter_direct(const T *__restrict__ pos, int64_t ps, const T *__restrict__ val,
                                                      int64_t vs, int64_t n, int ncomp, Grid g, int degree,
                                                      double *__restrict__ field,
                                                      unsigned long long *__restrict__ accum,
                                                      const double *__restrict__ scale_slot) {
  const double scale = MODE == MODE_FIXED ? scale_slot[0] : 1.0;
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    bool live = false;
    for (int c = 0; c < ncomp; ++c) live |= (double)val[c * vs + m] != 0.0;
    if (!live) continue;  // a dead marker: its position is not read
    int64_t node[3][4];
    double w[3][4];
    int size[3] = {1, 1, 1};
    node[1][0] = node[2][0] = 0;
    w[1][0] = w[2][0] = 1.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const double x = (double)pos[d * ps + m];
      // the 1D linear deposit divides by the spacing (scatter_linear_1d_pyccel.py:47)
      const double rel = (KIND == LINEAR && NDIM == 1) ? (x - g.lower[d]) / g.cell[d] : (x - g.lower[d]) * g.inverse[d];
      size[d] = axis_stencil<KIND>(rel, g.num[d], g.code[d], degree, node[d], w[d]);
    }
    for (int a = 0; a < size[0]; ++a)
      for (int b = 0; b < size[1]; ++b) {
        const double wab = NDIM >= 2 ? w[0][a] * w[1][b] : w[0][a];
        for (int e = 0; e < size[2]; ++e) {
          const double weight = NDIM == 3 ? wab * w[2][e] : wab;
          const int64_t at = (node[0][a] * g.stride[0] + node[1][b] * g.stride[1] + node[2][e] * g.stride[2]) * ncomp;
          for (int c = 0; c < ncomp; ++c) {
            const double v = (double)val[c * vs + m];
            if (KIND == LINEAR && NDIM == 1 && v == 0.0) continue;  // per-component kernel skips zeros
            const double contribution = weight * v;
            if (MODE == MODE_ATOMIC)
              unsafeAtomicAdd(field + at + c, contribution);
            else
              atomicAdd(accum + at + c, to_fixed(contribution * scale));
          }
        }
      }
  }
}

template <class T, int KIND, int NDIM>
__global__ void __launch_bounds__(256) gather_direct(const T *__restrict__ pos, int64_t ps, int64_t n, int ncomp, Grid g,
                                                     int degree, const double *__restrict__ field, T *__restrict__ res,
                                                     int64_t rs) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    int64_t node[3][4];
    double w[3][4];
    int size[3] = {1, 1, 1};
    node[1][0] = node[2][0] = 0;
    w[1][0] = w[2][0] = 1.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const double rel = ((double)pos[d * ps + m] - g.lower[d]) * g.inverse[d];
      size[d] = axis_stencil<KIND>(rel, g.num[d], g.code[d], degree, node[d], w[d]);
    }
    for (int c = 0; c < ncomp; ++c) {
      double total = 0.0;
      for (int a = 0; a < size[0]; ++a)
        for (int b = 0; b < size[1]; ++b) {
          const double wab = NDIM >= 2 ? w[0][a] * w[1][b] : w[0][a];
          for (int e = 0; e < size[2]; ++e) {
            const double weight = NDIM == 3 ? wab * w[2][e] : wab;
            total += weight *
                     field[(node[0][a] * g.stride[0] + node[1][b] * g.stride[1] + node[2][e] * g.stride[2]) * ncomp + c];
          }
        }
      res[c * rs + m] = (T)total;
    }
  }
}

// out[(d * ncomp + c) * os + m] = d/dx_d of the interpolant of component c at marker m.
template <class T, int KIND, int NDIM>
__global__ void __launch_bounds__(256) gather_derivative_direct(const T *__restrict__ pos, int64_t ps, int64_t n,
                                                                int ncomp, Grid g, int degree,
                                                                const double *__restrict__ field, T *__restrict__ out,
                                                                int64_t os) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    int64_t node[3][4];
    double w[3][4], dw[3][4];
    int size[3] = {1, 1, 1};
    node[1][0] = node[2][0] = 0;
    w[1][0] = w[2][0] = 1.0;
    dw[1][0] = dw[2][0] = 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const double rel = ((double)pos[d * ps + m] - g.lower[d]) * g.inverse[d];
      size[d] = axis_derivative<KIND>(rel, g.num[d], g.code[d], degree, node[d], w[d], dw[d]);
    }
    for (int c = 0; c < ncomp; ++c) {
      double total[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < size[0]; ++a)
        for (int b = 0; b < size[1]; ++b)
          for (int e = 0; e < size[2]; ++e) {
            const double f =
                field[(node[0][a] * g.stride[0] + node[1][b] * g.stride[1] + node[2][e] * g.stride[2]) * ncomp + c];
            total[0] += dw[0][a] * w[1][b] * w[2][e] * f;
            if (NDIM >= 2) total[1] += w[0][a] * dw[1][b] * w[2][e] * f;
            if (NDIM == 3) total[2] += w[0][a] * w[1][b] * dw[2][e] * f;
          }
#pragma unroll
      for (int d = 0; d < NDIM; ++d) out[(d * ncomp + c) * os + m] = (T)(total[d] * g.inverse[d]);
    }
  }
}

template <class T, int KIND, int NDIM>
static void launch_scatter(int mode, const void *pos, int64_t ps, const void *val, int64_t vs, int64_t n, int ncomp,
                           const Grid &g, int degree, double *field, unsigned long long *accum, const double *slot,
                           hipStream_t stream) {
  const int threads = 256;
  const int blocks = blocks_for(n, threads, 1 << 16);
  if (mode == MODE_ATOMIC)
    hipLaunchKernelGGL((scatter_direct<T, KIND, NDIM, MODE_ATOMIC>), dim3(blocks), dim3(threads), 0, stream,
                       (const T *)pos, ps, (const T *)val, vs, n, ncomp, g, degree, field, accum, slot);
  else
    hipLaunchKernelGGL((scatter_direct<T, KIND, NDIM, MODE_FIXED>), dim3(blocks), dim3(threads), 0, stream,
                       (const T *)pos, ps, (const T *)val, vs, n, ncomp, g, degree, field, accum, slot);
}

template <class T, int KIND>
static void dispatch_scatter_ndim(int ndim, int mode, const void *pos, int64_t ps, const void *val, int64_t vs,
                                  int64_t n, int ncomp, const Grid &g, int degree, double *field,
                                  unsigned long long *accum, const double *slot, hipStream_t stream) {
  if (ndim == 1) launch_scatter<T, KIND, 1>(mode, pos, ps, val, vs, n, ncomp, g, degree, field, accum, slot, stream);
  if (ndim == 2) launch_scatter<T, KIND, 2>(mode, pos, ps, val, vs, n, ncomp, g, degree, field, accum, slot, stream);
  if (ndim == 3) launch_scatter<T, KIND, 3>(mode, pos, ps, val, vs, n, ncomp, g, degree, field, accum, slot, stream);
}

template <class T>
static void dispatch_scatter(int kind, int ndim, int mode, const void *pos, int64_t ps, const void *val, int64_t vs,
                             int64_t n, int ncomp, const Grid &g, int degree, double *field, unsigned long long *accum,
                             const double *slot, hipStream_t stream) {
  if (kind == LINEAR)
    dispatch_scatter_ndim<T, LINEAR>(ndim, mode, pos, ps, val, vs, n, ncomp, g, degree, field, accum, slot, stream);
  else if (kind == QUADRATIC)
    dispatch_scatter_ndim<T, QUADRATIC>(ndim, mode, pos, ps, val, vs, n, ncomp, g, degree, field, accum, slot, stream);
  else
    dispatch_scatter_ndim<T, SPLINE>(ndim, mode, pos, ps, val, vs, n, ncomp, g, degree, field, accum, slot, stream);
}

template <class T, int KIND>
static void dispatch_gather_ndim(int ndim, const void *pos, int64_t ps, int64_t n, int ncomp, const Grid &g, int degree,
                                 const double *field, void *res, int64_t rs, hipStream_t stream) {
  const int threads = 256;
  const int blocks = blocks_for(n, threads, 1 << 16);
  if (ndim == 1)
    hipLaunchKernelGGL((gather_direct<T, KIND, 1>), dim3(blocks), dim3(threads), 0, stream, (const T *)pos, ps, n,
                       ncomp, g, degree, field, (T *)res, rs);
  if (ndim == 2)
    hipLaunchKernelGGL((gather_direct<T, KIND, 2>), dim3(blocks), dim3(threads), 0, stream, (const T *)pos, ps, n,
                       ncomp, g, degree, field, (T *)res, rs);
  if (ndim == 3)
    hipLaunchKernelGGL((gather_direct<T, KIND, 3>), dim3(blocks), dim3(threads), 0, stream, (const T *)pos, ps, n,
                       ncomp, g, degree, field, (T *)res, rs);
}

template <class T>
static void dispatch_gather(int kind, int ndim, const void *pos, int64_t ps, int64_t n, int ncomp, const Grid &g,
                            int degree, const double *field, void *res, int64_t rs, hipStream_t stream) {
  if (kind == LINEAR)
    dispatch_gather_ndim<T, LINEAR>(ndim, pos, ps, n, ncomp, g, degree, field, res, rs, stream);
  else if (kind == QUADRATIC)
    dispatch_gather_ndim<T, QUADRATIC>(ndim, pos, ps, n, ncomp, g, degree, field, res, rs, stream);
  else
    dispatch_gather_ndim<T, SPLINE>(ndim, pos, ps, n, ncomp, g, degree, field, res, rs, stream);
}

int check_common(int kind, int degree, int dtype, int64_t n, int64_t ncomp) {
  if (kind < LINEAR || kind > SPLINE) {
    set_error("kind must be 0 (linear), 1 (quadratic) or 2 (spline), got %d", kind);
    return MPF_ERR_ARGUMENT;
  }
  if (kind == SPLINE && (degree < 0 || degree > 3)) {
    set_error("spline degree must be 0 to 3, got %d", degree);
    return MPF_ERR_ARGUMENT;
  }
  if (dtype != F32 && dtype != F64) {
    set_error("dtype must be MPF_F32 or MPF_F64, got %d", dtype);
    return MPF_ERR_ARGUMENT;
  }
  if (n < 0 || ncomp < 1 || ncomp > 64) {
    set_error("need n >= 0 and 1 <= ncomp <= 64, got n=%lld ncomp=%lld", (long long)n, (long long)ncomp);
    return MPF_ERR_ARGUMENT;
  }
  return MPF_OK;
}

int kind_degree(int kind, int degree) { return kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree); }

}  // namespace mpf

using namespace mpf;

extern "C" int mpf_scatter(int kind, int degree, int ndim, const void *positions, int64_t pos_stride,
                           const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp, double *field,
                           const double *lower, const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                           int mode, int scale_exp, mpf_context *ctx, mpf_stream stream_) {
  int rc = check_common(kind, degree, dtype, n, ncomp);
  if (rc) return rc;
  if (mode != MODE_ATOMIC && mode != MODE_FIXED) {
    set_error("mode must be MPF_MODE_ATOMIC or MPF_MODE_FIXED, got %d", mode);
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  rc = make_grid(g, ndim, kind, kind_degree(kind, degree), lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  // large fixed-point deposits on shapes the tiled kernels take: binned by stencil cell
  // inside the call and deposited tile by tile (the same integers, bit for bit)
  if (mode == MODE_FIXED && scale_exp == MPF_SCALE_AUTO && binned_route(kind, degree, ndim, ncomp, g, n))
    return scatter_via_bins(kind, degree, ndim, positions, pos_stride, values, val_stride, dtype, n, ncomp, field,
                            lower, cell_size, num_nodes, code, ctx, stream);
  const int64_t count = g.extent[0] * (ndim > 1 ? g.extent[1] : 1) * (ndim > 2 ? g.extent[2] : 1) * ncomp;
  unsigned long long *accum = nullptr;
  const double *slot = nullptr;
  if (mode == MODE_FIXED) {
    rc = ensure(ctx->accum, (size_t)count * 8, true, stream);
    if (rc) return rc;
    rc = prepare_scale(ctx, values, val_stride, dtype, ncomp, n, 1.0, scale_exp, stream);
    if (rc) return rc;
    accum = (unsigned long long *)ctx->accum.ptr;
    slot = scale_slot(ctx);
  }
  {
    ProfileScope prof("scatter_direct", stream);
    if (dtype == F32)
      dispatch_scatter<float>(kind, ndim, mode, positions, pos_stride, values, val_stride, n, (int)ncomp, g,
                              kind_degree(kind, degree), field, accum, slot, stream);
    else
      dispatch_scatter<double>(kind, ndim, mode, positions, pos_stride, values, val_stride, n, (int)ncomp, g,
                               kind_degree(kind, degree), field, accum, slot, stream);
  }
  MPF_LAUNCH_CHECK();
  if (mode == MODE_FIXED) return finalize_fixed(ctx, field, count, stream);
  return MPF_OK;
}

// ---- the deposit that allocates nothing ------------------------------------------------
//
// mpf_scatter_noalloc keeps its fixed-point accumulator and scale slot in `scratch`, a
// file-scope device array (one copy per GPU, held until the process ends): [0, 256) the
// max |value| bits and the scale slot, then MPF_NOALLOC_MAX_VALUES int64 accumulator words.
// Host code never passes the array's name to a kernel (on the host it is a stand-in, not a
// device address): the device address comes from hipGetSymbolAddress, once per device.

namespace mpf {

constexpr size_t NOALLOC_HEAD = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_HEAD + (size_t)MPF_NOALLOC_MAX_VALUES * 8;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mutex;
static std::map<int, unsigned char *> g_scratch_ptr;  // device -> the address of its copy

// The device address of this device's scratch, looked up on the first call on the device.
static int noalloc_scratch(unsigned char **out) {
  int dev = 0;
  MPF_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void *ptr = nullptr;
    MPF_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    it = g_scratch_ptr.emplace(dev, (unsigned char *)ptr).first;
  }
  *out = it->second;
  return MPF_OK;
}

}  // namespace mpf

extern "C" int mpf_scatter_noalloc(int kind, int degree, int ndim, const void *positions, int64_t pos_stride,
                                   const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp,
                                   double *field, const double *lower, const double *cell_size,
                                   const int64_t *num_nodes, const int64_t *code, int mode, int scale_exp,
                                   mpf_stream stream_) {
  int rc = check_common(kind, degree, dtype, n, ncomp);
  if (rc) return rc;
  if (mode != MODE_ATOMIC && mode != MODE_FIXED) {
    set_error("mode must be MPF_MODE_ATOMIC or MPF_MODE_FIXED, got %d", mode);
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  rc = make_grid(g, ndim, kind, kind_degree(kind, degree), lower, cell_size, num_nodes, code);
  if (rc) return rc;
  const int64_t count = g.extent[0] * (ndim > 1 ? g.extent[1] : 1) * (ndim > 2 ? g.extent[2] : 1) * ncomp;
  if (mode == MODE_FIXED && count > (int64_t)MPF_NOALLOC_MAX_VALUES) {
    set_error("scatter_noalloc: a fixed-point deposit takes at most %lld field values (MPF_NOALLOC_MAX_VALUES), "
              "this field has %lld",
              (long long)MPF_NOALLOC_MAX_VALUES, (long long)count);
    return MPF_ERR_ARGUMENT;
  }
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  unsigned long long *accum = nullptr;
  double *slot = nullptr;
  if (mode == MODE_FIXED) {
    unsigned char *base = nullptr;
    rc = noalloc_scratch(&base);
    if (rc) return rc;
    unsigned long long *maxbits = (unsigned long long *)base;
    slot = (double *)(base + 64);
    accum = (unsigned long long *)(base + NOALLOC_HEAD);
    // the array keeps the last call's contents: the accumulator starts at zero
    rc = zero_async(accum, (size_t)count * 8, stream);
    if (rc) return rc;
    rc = scale_into(values, val_stride, dtype, ncomp, n, 1.0, scale_exp, maxbits, slot, stream);
    if (rc) return rc;
  }
  if (dtype == F32)
    dispatch_scatter<float>(kind, ndim, mode, positions, pos_stride, values, val_stride, n, (int)ncomp, g,
                            kind_degree(kind, degree), field, accum, slot, stream);
  else
    dispatch_scatter<double>(kind, ndim, mode, positions, pos_stride, values, val_stride, n, (int)ncomp, g,
                             kind_degree(kind, degree), field, accum, slot, stream);
  MPF_LAUNCH_CHECK();
  if (mode == MODE_FIXED) return finalize_into(field, accum, count, slot, stream);
  return MPF_OK;
}

extern "C" int mpf_gather(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, int dtype,
                          int64_t n, const double *field, int64_t ncomp, void *result, int64_t result_stride,
                          const double *lower, const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                          mpf_context *ctx, mpf_stream stream_) {
  int rc = check_common(kind, degree, dtype, n, ncomp);
  if (rc) return rc;
  Grid g;
  rc = make_grid(g, ndim, kind, kind_degree(kind, degree), lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  // large gathers on shapes the tiled kernels take: binned inside the call, gathered tile by
  // tile, and the results put back in the input order
  if (binned_route(kind, degree, ndim, ncomp, g, n)) {
    if (!ctx) ctx = default_context(stream);
    if (!ctx) return MPF_ERR_HIP;
    return gather_via_bins(kind, degree, ndim, positions, pos_stride, dtype, n, field, ncomp, result, result_stride,
                           lower, cell_size, num_nodes, code, ctx, stream);
  }
  ProfileScope prof("gather_direct", stream);
  if (dtype == F32)
    dispatch_gather<float>(kind, ndim, positions, pos_stride, n, (int)ncomp, g, kind_degree(kind, degree), field,
                           result, result_stride, stream);
  else
    dispatch_gather<double>(kind, ndim, positions, pos_stride, n, (int)ncomp, g, kind_degree(kind, degree), field,
                            result, result_stride, stream);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

template <class T, int KIND>
static void dispatch_derivative_ndim(int ndim, const void *pos, int64_t ps, int64_t n, int ncomp, const Grid &g,
                                     int degree, const double *field, void *out, int64_t os, hipStream_t stream) {
  const int threads = 256, blocks = blocks_for(n, threads, 1 << 16);
  if (ndim == 1)
    hipLaunchKernelGGL((gather_derivative_direct<T, KIND, 1>), dim3(blocks), dim3(threads), 0, stream, (const T *)pos,
                       ps, n, ncomp, g, degree, field, (T *)out, os);
  if (ndim == 2)
    hipLaunchKernelGGL((gather_derivative_direct<T, KIND, 2>), dim3(blocks), dim3(threads), 0, stream, (const T *)pos,
                       ps, n, ncomp, g, degree, field, (T *)out, os);
  if (ndim == 3)
    hipLaunchKernelGGL((gather_derivative_direct<T, KIND, 3>), dim3(blocks), dim3(threads), 0, stream, (const T *)pos,
                       ps, n, ncomp, g, degree, field, (T *)out, os);
}

extern "C" int mpf_gather_derivative(int kind, int degree, int ndim, const void *positions, int64_t pos_stride,
                                     int dtype, int64_t n, const double *field, int64_t ncomp, void *out,
                                     int64_t out_stride, const double *lower, const double *cell_size,
                                     const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                                     mpf_stream stream_) {
  (void)ctx;
  int rc = check_common(kind, degree, dtype, n, ncomp);
  if (rc) return rc;
  for (int d = 0; d < ndim && d < 3; ++d)
    if (code[d] == 2) {
      set_error("the derivative gather does not take clamped axes");
      return MPF_ERR_UNSUPPORTED;
    }
  Grid g;
  rc = make_grid(g, ndim, kind, kind_degree(kind, degree), lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  const int deg = kind_degree(kind, degree);
#define DER(T)                                                                                                    \
  if (kind == LINEAR)                                                                                             \
    dispatch_derivative_ndim<T, LINEAR>(ndim, positions, pos_stride, n, (int)ncomp, g, deg, field, out, out_stride, \
                                        stream);                                                                  \
  else if (kind == QUADRATIC)                                                                                     \
    dispatch_derivative_ndim<T, QUADRATIC>(ndim, positions, pos_stride, n, (int)ncomp, g, deg, field, out,        \
                                           out_stride, stream);                                                   \
  else                                                                                                            \
    dispatch_derivative_ndim<T, SPLINE>(ndim, positions, pos_stride, n, (int)ncomp, g, deg, field, out, out_stride, \
                                        stream);
  if (dtype == F32) {
    DER(float)
  } else {
    DER(double)
  }
#undef DER
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

// ---- the catalogue's entry points, argument for argument ----------------------------

#define GEOM1 MPF_GEOM_1D
#define GEOM2 MPF_GEOM_2D
#define GEOM3 MPF_GEOM_3D
#define G1 const double lo[1] = {lower_x}; const double cs[1] = {cell_size_x}; const int64_t nn[1] = {num_x}; \
  const int64_t pc[1] = {periodic_x};
#define G2 const double lo[2] = {lower_x, lower_y}; const double cs[2] = {cell_size_x, cell_size_y}; \
  const int64_t nn[2] = {num_x, num_y}; const int64_t pc[2] = {periodic_x, periodic_y};
#define G3 const double lo[3] = {lower_x, lower_y, lower_z}; \
  const double cs[3] = {cell_size_x, cell_size_y, cell_size_z}; const int64_t nn[3] = {num_x, num_y, num_z}; \
  const int64_t pc[3] = {periodic_x, periodic_y, periodic_z};

#define NODE_ENTRY(KINDNAME, KIND, D)                                                                              \
  extern "C" int mpf_scatter_##KINDNAME##_##D##d(MPF_SCATTER_IN, GEOM##D, int mode, MPF_TAIL) {           \
    G##D return mpf_scatter(KIND, 0, D, positions, pos_stride, values, val_stride, dtype, n, ncomp, field, lo, cs, \
                            nn, pc, mode, MPF_SCALE_AUTO, ctx, stream);                                           \
  }                                                                                                               \
  extern "C" int mpf_gather_##KINDNAME##_##D##d(MPF_GATHER_IN, GEOM##D, MPF_TAIL) {                        \
    G##D return mpf_gather(KIND, 0, D, positions, pos_stride, dtype, n, field, ncomp, result, result_stride, lo,   \
                           cs, nn, pc, ctx, stream);                                                              \
  }

NODE_ENTRY(linear, MPF_KIND_LINEAR, 1)
NODE_ENTRY(linear, MPF_KIND_LINEAR, 2)
NODE_ENTRY(linear, MPF_KIND_LINEAR, 3)
NODE_ENTRY(quadratic, MPF_KIND_QUADRATIC, 1)
NODE_ENTRY(quadratic, MPF_KIND_QUADRATIC, 2)
NODE_ENTRY(quadratic, MPF_KIND_QUADRATIC, 3)

#define SPLINE_ENTRY(D)                                                                                             \
  extern "C" int mpf_scatter_spline_##D##d(MPF_SCATTER_IN, GEOM##D, int degree, int mode, MPF_TAIL) {      \
    G##D return mpf_scatter(MPF_KIND_SPLINE, degree, D, positions, pos_stride, values, val_stride, dtype, n, ncomp, \
                            field, lo, cs, nn, pc, mode, MPF_SCALE_AUTO, ctx, stream);                             \
  }                                                                                                                \
  extern "C" int mpf_gather_spline_##D##d(MPF_GATHER_IN, GEOM##D, int degree, MPF_TAIL) {                   \
    G##D return mpf_gather(MPF_KIND_SPLINE, degree, D, positions, pos_stride, dtype, n, field, ncomp, result,       \
                           result_stride, lo, cs, nn, pc, ctx, stream);                                            \
  }

SPLINE_ENTRY(1)
SPLINE_ENTRY(2)
SPLINE_ENTRY(3)
