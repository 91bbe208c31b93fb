// mpf_common.hip -- library plumbing: errors, contexts, scratch, the fixed-point scale.
#include <math.h>
#include <stdarg.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mpf_internal.h"

namespace mpf {

static thread_local char g_error[512] = "";

// Clears device memory with a kernel rather than hipMemsetAsync, so that a caller may
// capture the library's work in a HIP graph.
__global__ void zero_kernel(unsigned long long *__restrict__ p, int64_t words) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

int zero_async(void *p, size_t bytes, hipStream_t stream) {
  const int64_t words = (int64_t)((bytes + 7) / 8);
  if (words == 0) return MPF_OK;
  hipLaunchKernelGGL(zero_kernel, dim3(blocks_for(words, 256, 1 << 14)), dim3(256), 0, stream,
                     (unsigned long long *)p, words);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

int fail_hip(hipError_t e, const char *what) {
  set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return MPF_ERR_HIP;
}

int ensure(Buffer &b, size_t bytes, bool zero, hipStream_t stream) {
  if (bytes <= b.bytes && b.ptr) return MPF_OK;
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (b.ptr) {
    // the old buffer may still be read by queued work on this stream
    MPF_CHECK(hipStreamSynchronize(stream));
    MPF_CHECK(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
  }
  MPF_CHECK(hipMalloc(&b.ptr, want));
  if (zero) {
    int rc = zero_async(b.ptr, want, stream);
    if (rc) return rc;
  }
  b.bytes = want;
  return MPF_OK;
}

static std::mutex g_ctx_mutex;
static std::map<std::pair<int, hipStream_t>, mpf_context *> g_default;

// One default context per (device, stream): calls on different streams never share scratch,
// and a context's buffers are reallocated only after its own stream has drained.
mpf_context *default_context(hipStream_t stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  mpf_context *&ctx = g_default[{dev, stream}];
  if (!ctx) ctx = mpf_context_create(dev);
  return ctx;
}

int make_grid(Grid &g, int ndim, int kind, int degree, const double *lower, const double *cell, const int64_t *num,
              const int64_t *code) {
  if (ndim < 1 || ndim > 3) {
    set_error("ndim must be 1, 2 or 3, got %d", ndim);
    return MPF_ERR_ARGUMENT;
  }
  g.ndim = ndim;
  for (int d = 0; d < 3; ++d) {
    const bool on = d < ndim;
    g.lower[d] = on ? lower[d] : 0.0;
    g.cell[d] = on ? cell[d] : 1.0;
    g.inverse[d] = 1.0 / g.cell[d];
    g.num[d] = on ? num[d] : 1;
    g.code[d] = on ? (int)code[d] : 0;
    if (on && (code[d] < 0 || code[d] > 2)) {
      set_error("boundary code of axis %d must be 0, 1 or 2, got %lld", d, (long long)code[d]);
      return MPF_ERR_ARGUMENT;
    }
    if (on && code[d] == 2 && kind != SPLINE && kind != 100) {
      set_error("a clamped axis (code 2) needs the spline kernels");
      return MPF_ERR_ARGUMENT;
    }
    if (on && num[d] < 2) {
      set_error("axis %d needs at least 2 nodes, got %lld", d, (long long)num[d]);
      return MPF_ERR_ARGUMENT;
    }
    g.extent[d] = g.num[d] + ((on && g.code[d] == 2) ? (degree - 1) : 0);
  }
  g.stride[2] = 1;
  g.stride[1] = g.extent[2];
  g.stride[0] = g.extent[1] * g.extent[2];
  if (ndim == 1) {
    g.stride[0] = 1;
  } else if (ndim == 2) {
    g.stride[0] = g.extent[1];
    g.stride[1] = 1;
  }
  return MPF_OK;
}

// ---------------------------------------------------------------------------------------
// The fixed-point scale, on the device.

template <class T>
__global__ void max_abs_kernel(const T *__restrict__ v, int64_t vs, int64_t rows, int64_t n,
                               unsigned long long *__restrict__ out) {
  double m = 0.0;
  // row by row, four loads of the grid-stride walk in flight per trip (no index division)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x, t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t r = 0; r < rows; ++r) {
    const T *row = v + r * vs;
    int64_t c = t;
    for (; c + 3 * stride < n; c += 4 * stride) {
      const double a0 = fabs((double)row[c]), a1 = fabs((double)row[c + stride]);
      const double a2 = fabs((double)row[c + 2 * stride]), a3 = fabs((double)row[c + 3 * stride]);
      // NaN compares false and is left out of the bound
      if (a0 > m) m = a0;
      if (a1 > m) m = a1;
      if (a2 > m) m = a2;
      if (a3 > m) m = a3;
    }
    for (; c < n; c += stride) {
      const double a = fabs((double)row[c]);
      if (a > m) m = a;
    }
  }
  // wavefront reduction, then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

__global__ void scale_kernel(const unsigned long long *__restrict__ maxbits, double count, double factor, int fixed_exp,
                             double *__restrict__ slot) {
  // the largest |value|, 0 when none was measured
  const double m = __longlong_as_double((long long)maxbits[0]);
  int k;
  if (fixed_exp != MPF_SCALE_AUTO) {
    k = fixed_exp;
  } else {
    const double bound = m * count * factor;
    if (!(bound > 0.0) || isinf(bound)) {
      k = 0;
    } else {
      // every node sum below 2^61 quanta (bound < 2^(ilogb + 1)), and one marker's value
      // below 2^49 quanta so that the float64 register sums stay exact
      k = 61 - (ilogb(bound) + 1);
      const int single = 49 - (ilogb(m) + 1);
      if (single < k) k = single;
      if (k > 900) k = 900;
      if (k < -900) k = -900;
    }
  }
  slot[0] = ldexp(1.0, k);
  slot[1] = ldexp(1.0, -k);
  // markers per float64 register pass: |sum| <= chunk * max|value| * 2^k < 2^50
  double chunk = 1.0e18;
  if (m > 0.0 && !isinf(m)) {
    chunk = floor(ldexp(1.0, 50 - k) / m);
    if (chunk < 1.0) chunk = 1.0;
    if (chunk > 1.0e18) chunk = 1.0e18;
  }
  slot[2] = chunk;
}

double *scale_slot(mpf_context *ctx) { return (double *)ctx->scalars.ptr + 1; }

int prepare_scale(mpf_context *ctx, const void *values, int64_t vstride, int dtype, int64_t rows, int64_t n,
                  double bound_factor, int scale_exp, hipStream_t stream) {
  int rc = ensure(ctx->scalars, 64, true, stream);
  if (rc) return rc;
  return scale_into(values, vstride, dtype, rows, n, bound_factor, scale_exp, (unsigned long long *)ctx->scalars.ptr,
                    scale_slot(ctx), stream);
}

int scale_into(const void *values, int64_t vstride, int dtype, int64_t rows, int64_t n, double bound_factor,
               int scale_exp, unsigned long long *maxbits, double *slot, hipStream_t stream) {
  int rc;
  if (scale_exp == MPF_SCALE_AUTO && values) {  // values == nullptr: the caller filled the bound slot
    rc = zero_async(maxbits, 8, stream);
    if (rc) return rc;
    if (n > 0) {
      const int threads = 256;
      const int blocks = blocks_for(n, threads, 4096);
      if (dtype == F32)
        hipLaunchKernelGGL(max_abs_kernel<float>, dim3(blocks), dim3(threads), 0, stream, (const float *)values,
                           vstride, rows, n, maxbits);
      else
        hipLaunchKernelGGL(max_abs_kernel<double>, dim3(blocks), dim3(threads), 0, stream, (const double *)values,
                           vstride, rows, n, maxbits);
      MPF_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(scale_kernel, dim3(1), dim3(1), 0, stream, maxbits, (double)n, bound_factor, scale_exp, slot);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

__global__ void finalize_kernel(double *__restrict__ field, unsigned long long *__restrict__ accum, int64_t count,
                                const double *__restrict__ slot) {
  const double inv = slot[1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const long long q = (long long)accum[i];
    if (q != 0) {
      field[i] += (double)q * inv;
      accum[i] = 0ull;
    }
  }
}

int finalize_into(double *field, unsigned long long *accum, int64_t count, const double *slot, hipStream_t stream) {
  if (count == 0) return MPF_OK;
  const int threads = 256;
  hipLaunchKernelGGL(finalize_kernel, dim3(blocks_for(count, threads, 1 << 16)), dim3(threads), 0, stream, field,
                     accum, count, slot);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

int finalize_fixed_with(mpf_context *ctx, double *field, int64_t count, const double *slot, hipStream_t stream) {
  return finalize_into(field, (unsigned long long *)ctx->accum.ptr, count, slot, stream);
}

int finalize_fixed(mpf_context *ctx, double *field, int64_t count, hipStream_t stream) {
  return finalize_into(field, (unsigned long long *)ctx->accum.ptr, count, scale_slot(ctx), stream);
}

// ---- kernel timing ---------------------------------------------------------------------

struct EventPair {
  hipEvent_t start, stop;
};

static std::mutex g_prof_mutex;
static bool g_prof_on = false;
static std::map<std::string, std::vector<EventPair>> g_prof;

ProfileScope::ProfileScope(const char *name_, hipStream_t stream_) : name(name_), stream(stream_), slot(-1) {
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  if (!g_prof_on) return;
  EventPair e;
  if (hipEventCreate(&e.start) != hipSuccess || hipEventCreate(&e.stop) != hipSuccess) return;
  (void)hipEventRecord(e.start, stream);
  auto &v = g_prof[name];
  v.push_back(e);
  slot = (int)v.size() - 1;
}

ProfileScope::~ProfileScope() {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  (void)hipEventRecord(g_prof[name][slot].stop, stream);
}

}  // namespace mpf

extern "C" {

/* Timing of the library's kernels (see include/mpf_hip.h). */
int mpf_profile_enable(int on) {
  std::lock_guard<std::mutex> lock(mpf::g_prof_mutex);
  mpf::g_prof_on = on != 0;
  return MPF_OK;
}

int mpf_profile_reset(void) {
  std::lock_guard<std::mutex> lock(mpf::g_prof_mutex);
  for (auto &kv : mpf::g_prof)
    for (auto &e : kv.second) {
      (void)hipEventDestroy(e.start);
      (void)hipEventDestroy(e.stop);
    }
  mpf::g_prof.clear();
  return MPF_OK;
}

int mpf_profile_query(const char *name, double *total_ms, int64_t *launches) {
  std::lock_guard<std::mutex> lock(mpf::g_prof_mutex);
  *total_ms = 0.0;
  *launches = 0;
  auto it = mpf::g_prof.find(name);
  if (it == mpf::g_prof.end()) return MPF_OK;
  for (auto &e : it->second) {
    if (hipEventSynchronize(e.stop) != hipSuccess) return mpf::fail_hip(hipErrorUnknown, "profile query");
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e.start, e.stop) != hipSuccess) continue;
    *total_ms += ms;
    *launches += 1;
  }
  return MPF_OK;
}


const char *mpf_last_error(void) { return mpf::g_error; }

int mpf_abi_version(void) { return 1; }

mpf_context *mpf_context_create(int device) {
  int old = 0;
  if (hipGetDevice(&old) != hipSuccess) {
    mpf::set_error("no HIP device");
    return nullptr;
  }
  if (device != old && hipSetDevice(device) != hipSuccess) {
    mpf::set_error("cannot select device %d", device);
    return nullptr;
  }
  mpf_context *ctx = new mpf_context();
  ctx->device = device;
  if (device != old) (void)hipSetDevice(old);
  return ctx;
}

void mpf_context_destroy(mpf_context *ctx) {
  if (!ctx) return;
  (void)hipDeviceSynchronize();
  mpf::Buffer *all[] = {&ctx->accum,   &ctx->scalars, &ctx->sort_keys, &ctx->sort_tmp,    &ctx->offsets,
                        &ctx->staging, &ctx->path_in, &ctx->bins,      &ctx->path_sorted, &ctx->rebin,
                        &ctx->perm_plan, &ctx->perm_build, &ctx->perm_stage};
  for (mpf::Buffer *b : all)
    if (b->ptr) (void)hipFree(b->ptr);
  delete ctx;
}

}  // extern "C"
