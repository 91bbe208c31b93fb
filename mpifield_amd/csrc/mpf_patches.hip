// mpf_patches.hip -- many refined-mesh patches in one persistent launch.
//
// The reference deposits into a patch with one Weighting.scatter_block call per patch
// (BlockGrid.from_geometry(lower, cell, num, periodic), _types.py:46-96; the caller loops,
// edges.py:21).  Here a plan holds every patch's geometry, kind and degree and a work list
// of (patch, chunk of markers) items; one persistent kernel's workgroups pull items from an
// atomic counter until the list is drained.  An item deposits its markers into an LDS copy
// of its patch as int64 fixed point and adds the non-zero nodes into the patch's int64
// accumulator; a finalize pass adds the accumulators into the float64 patch fields.  The
// sum is exact in integers, so the result does not depend on which workgroup took which
// item.  The launches take no host synchronisation and allocate nothing once the plan and
// the context's scratch exist, so a caller can capture them in a HIP graph.
#include <vector>

#include "mpf_internal.h"

struct mpf_patch_plan {
  int ndim = 0;
  int64_t n_patches = 0, n_items = 0, chunk = 0, field_values = 0, max_nodes = 0, markers = 0;
  void *patches = nullptr;   // device array of mpf_patch
  void *items = nullptr;     // device array of int64 triples (patch, first marker, count)
  unsigned int *counter = nullptr;
};

namespace mpf {

struct Item {
  int64_t patch, first, count;
};

// The dynamic LDS copy of a patch as an LDS pointer (address space 3).
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

// One axis of a patch weighting with int32 nodes: linear_axis, quadratic_axis and
// spline_axis's arithmetic with the kind and degree fixed (the rel wrap of a periodic axis
// or clamp of a bounded one, the cell choice, the weights, the node wrap or clamp).
template <int KIND, int P>
struct PatchStencil {
  static constexpr int S = KIND == LINEAR ? 2 : (KIND == QUADRATIC ? 3 : P + 1);
};

template <int KIND, int P>
__device__ __forceinline__ int patch_stencil(double rel, int num, int code, int *node, double *w) {
  double r = rel;
  if (KIND == LINEAR) {
    int l;
    if (code == 1) r = periodic_position(r, num);
    // the cast sees r itself for any finite position (and 0 for NaN: no index outside)
    const int i = (int)fmin(fmax(r, 0.0), (double)num);
    if (code == 1)
      l = i >= num ? num - 1 : i;
    else
      l = i > num - 2 ? num - 2 : i;
    node[0] = l;
    node[1] = (code == 1 && l + 1 == num) ? 0 : l + 1;
    const double d = r - (double)l;
    const double f = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
    w[0] = 1.0 - f;
    w[1] = f;
    return l;
  }
  if (code == 1) {
    if (!(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
  } else {
    if (r < 0.0) r = 0.0;
    if (r > (double)(num - 1)) r = (double)num - 1.0;
  }
  constexpr int Q = KIND == QUADRATIC ? 2 : P;
  const double rc = fmin(fmax(r, 0.0), (double)num);  // r itself for any finite position, 0 for NaN
  const int middle = (Q % 2 == 0) ? (int)(rc + 0.5) : (int)rc;
  const double o = r - (double)middle;
  if (Q == 0) {
    w[0] = 1.0;
  } else if (Q == 1) {
    w[0] = 1.0 - o;
    w[1] = o;
  } else if (Q == 2) {
    w[0] = 0.5 * (0.5 - o) * (0.5 - o);
    w[1] = 0.75 - o * o;
    w[2] = 0.5 * (0.5 + o) * (0.5 + o);
  } else {
    const double cm = 1.0 - o, sq = o * o, cu = sq * o;
    w[0] = cm * cm * cm / 6.0;
    w[1] = (3.0 * cu - 6.0 * sq + 4.0) / 6.0;
    w[2] = (-3.0 * cu + 3.0 * sq + 3.0 * o + 1.0) / 6.0;
    w[3] = cu / 6.0;
  }
#pragma unroll
  for (int a = 0; a <= Q; ++a) {
    const int k = middle - Q / 2 + a;
    if (code == 1)
      node[a] = k < 0 ? k + num : (k >= num ? k - num : k);
    else
      node[a] = k < 0 ? 0 : (k > num - 1 ? num - 1 : k);
  }
  return middle;
}

// The node offsets of one marker of patch p (each axis's stencil, row-major strides), and
// its weights; false for a dead marker.
template <int NDIM, int KIND, int P, class T>
__device__ __forceinline__ int patch_marker(const mpf_patch &p, const T *__restrict__ pos, int64_t ps, int64_t m,
                                            int (&node)[3][4], double (&w)[3][4]) {
  int key = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const double rel = ((double)pos[d * ps + m] - p.lower[d]) * (1.0 / p.cell_size[d]);
    const int anchor = patch_stencil<KIND, P>(rel, (int)p.num_nodes[d], (int)p.code[d], node[d], w[d]);
    key = key * ((int)p.num_nodes[d] + 1) + anchor;  // anchors lie in [0, num]
  }
  return key;
}

template <int NDIM, int KIND, int P, class T>
__device__ __forceinline__ void patch_scatter_item(const mpf_patch &p, const T *__restrict__ pos, int64_t ps,
                                                   const T *__restrict__ val, int64_t vs, int ncomp, int64_t first,
                                                   int64_t count, lds_u64 *lds, double scale) {
  constexpr int S = PatchStencil<KIND, P>::S;
  const int s1 = NDIM == 3 ? (int)p.num_nodes[2] : 1;
  const int s0 = NDIM == 3 ? (int)(p.num_nodes[1] * p.num_nodes[2]) : (NDIM == 2 ? (int)p.num_nodes[1] : 1);
  for (int64_t m = first + threadIdx.x; m < first + count; m += blockDim.x) {
    bool live = false;
    for (int c = 0; c < ncomp; ++c) live |= (double)val[c * vs + m] != 0.0;
    if (!live) continue;
    int node[3][4];
    double w[3][4];
    patch_marker<NDIM, KIND, P, T>(p, pos, ps, m, node, w);
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
      for (int b = 0; b < (NDIM >= 2 ? S : 1); ++b) {
        const double wab = NDIM >= 2 ? w[0][a] * w[NDIM >= 2 ? 1 : 0][b] : w[0][a];
#pragma unroll
        for (int e = 0; e < (NDIM == 3 ? S : 1); ++e) {
          const double weight = NDIM == 3 ? wab * w[NDIM - 1][e] : wab;
          const int at = (node[0][a] * s0 + (NDIM >= 2 ? node[NDIM >= 2 ? 1 : 0][b] * s1 : 0) +
                          (NDIM == 3 ? node[NDIM - 1][e] : 0)) * ncomp;
          // the automatic scale keeps one marker's value below 2^49 quanta: to_fixed_small
          for (int c = 0; c < ncomp; ++c)
            __hip_atomic_fetch_add(lds + at + c, to_fixed_small(weight * (double)val[c * vs + m] * scale),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
  }
}

// The deposit of an item's markers of one component in runs: work-item t takes runs of
// RUN consecutive markers, run t of each pass of blockDim runs (a wavefront's loads stay
// near each other), and sums the contributions of consecutive markers with the same stencil
// (the same anchor cell) in float64 registers, exactly (the biased fma of
// odd_significand), adding them into the LDS copy when the anchor changes, after
// `per_pass` markers (the scale's bound on one register sum) and at the end of the run.  On
// markers kept sorted by cell within their patch, a run is mostly one cell: one LDS add
// per node and cell instead of one per marker.  Each contribution is rounded on its own, so
// the sums are the same integers whatever the runs.
template <int NDIM, int KIND, int P, class T>
__device__ __forceinline__ void patch_scatter_runs(const mpf_patch &p, const T *__restrict__ pos, int64_t ps,
                                                   const T *__restrict__ val, int64_t first, int64_t count,
                                                   lds_u64 *lds, double scale, int per_pass) {
  constexpr int S = PatchStencil<KIND, P>::S, RUN = 8;
  constexpr int SB = NDIM >= 2 ? S : 1, SE = NDIM == 3 ? S : 1;
  const int s1 = NDIM == 3 ? (int)p.num_nodes[2] : 1;
  const int s0 = NDIM == 3 ? (int)(p.num_nodes[1] * p.num_nodes[2]) : (NDIM == 2 ? (int)p.num_nodes[1] : 1);
  const int64_t end = first + count;
  for (int64_t lo = first + (int64_t)threadIdx.x * RUN; lo < end; lo += (int64_t)blockDim.x * RUN) {
    const int64_t hi = lo + RUN < end ? lo + RUN : end;
    double acc[S][SB][SE];
    int at[S][SB][SE];
    int key = -1, in_pass = 0;
    for (int64_t m = lo; m < hi; ++m) {
      const double v = (double)val[m];
      if (v == 0.0) continue;  // a dead marker deposits nothing
      int node[3][4];
      double w[3][4];
      const int k = patch_marker<NDIM, KIND, P, T>(p, pos, ps, m, node, w);
      if (k != key || in_pass >= per_pass) {
        if (key >= 0) {
#pragma unroll
          for (int a = 0; a < S; ++a)
#pragma unroll
            for (int b = 0; b < SB; ++b)
#pragma unroll
              for (int e = 0; e < SE; ++e)
                __hip_atomic_fetch_add(lds + at[a][b][e], (unsigned long long)biased_to_int(acc[a][b][e]),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        key = k;
        in_pass = 0;
#pragma unroll
        for (int a = 0; a < S; ++a)
#pragma unroll
          for (int b = 0; b < SB; ++b)
#pragma unroll
            for (int e = 0; e < SE; ++e) {
              acc[a][b][e] = BIAS;
              at[a][b][e] = node[0][a] * s0 + (NDIM >= 2 ? node[NDIM >= 2 ? 1 : 0][b] * s1 : 0) +
                            (NDIM == 3 ? node[NDIM - 1][e] : 0);
            }
      }
      const double u = odd_significand(v * scale);
#pragma unroll
      for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b = 0; b < SB; ++b) {
          const double wab = NDIM >= 2 ? w[0][a] * w[NDIM >= 2 ? 1 : 0][b] : w[0][a];
#pragma unroll
          for (int e = 0; e < SE; ++e)
            acc[a][b][e] = __builtin_fma(u, NDIM == 3 ? wab * w[NDIM - 1][e] : wab, acc[a][b][e]);
        }
      ++in_pass;
    }
    if (key >= 0) {
#pragma unroll
      for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b = 0; b < SB; ++b)
#pragma unroll
          for (int e = 0; e < SE; ++e)
            __hip_atomic_fetch_add(lds + at[a][b][e], (unsigned long long)biased_to_int(acc[a][b][e]),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
}

template <int NDIM, int KIND, int P, class T>
__device__ __forceinline__ void patch_gather_item(const mpf_patch &p, const T *__restrict__ pos, int64_t ps,
                                                  int ncomp, int64_t first, int64_t count,
                                                  const double *__restrict__ field, T *__restrict__ res, int64_t rs) {
  constexpr int S = PatchStencil<KIND, P>::S;
  const int s1 = NDIM == 3 ? (int)p.num_nodes[2] : 1;
  const int s0 = NDIM == 3 ? (int)(p.num_nodes[1] * p.num_nodes[2]) : (NDIM == 2 ? (int)p.num_nodes[1] : 1);
  for (int64_t m = first + threadIdx.x; m < first + count; m += blockDim.x) {
    int node[3][4];
    double w[3][4];
    patch_marker<NDIM, KIND, P, T>(p, pos, ps, m, node, w);
    for (int c = 0; c < ncomp; ++c) {
      double total = 0.0;
#pragma unroll
      for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b = 0; b < (NDIM >= 2 ? S : 1); ++b) {
          const double wab = NDIM >= 2 ? w[0][a] * w[NDIM >= 2 ? 1 : 0][b] : w[0][a];
#pragma unroll
          for (int e = 0; e < (NDIM == 3 ? S : 1); ++e) {
            const double weight = NDIM == 3 ? wab * w[NDIM - 1][e] : wab;
            total += weight * field[(node[0][a] * s0 + (NDIM >= 2 ? node[NDIM >= 2 ? 1 : 0][b] * s1 : 0) +
                                     (NDIM == 3 ? node[NDIM - 1][e] : 0)) * ncomp + c];
          }
        }
      res[c * rs + m] = (T)total;
    }
  }
}

// The item's kind and degree, fixed for the call F<NDIM, KIND, P>.
#define PATCH_DISPATCH(NDIM, CALL)                                   \
  switch (p.kind == LINEAR ? 10 : (p.kind == QUADRATIC ? 20 : p.degree)) { \
    case 10: CALL(NDIM, LINEAR, 1); break;                            \
    case 20: CALL(NDIM, QUADRATIC, 2); break;                         \
    case 0: CALL(NDIM, SPLINE, 0); break;                             \
    case 1: CALL(NDIM, SPLINE, 1); break;                             \
    case 2: CALL(NDIM, SPLINE, 2); break;                             \
    default: CALL(NDIM, SPLINE, 3); break;                            \
  }

__global__ void zero_counter(unsigned int *c) { *c = 0u; }

template <class T, int NDIM>
__global__ void __launch_bounds__(512) patches_scatter_kernel(const T *__restrict__ pos, int64_t ps,
                                                              const T *__restrict__ val, int64_t vs, int ncomp,
                                                              const mpf_patch *__restrict__ patches,
                                                              const Item *__restrict__ items, int64_t n_items,
                                                              unsigned int *__restrict__ counter,
                                                              unsigned long long *__restrict__ accum,
                                                              const double *__restrict__ slot, int lds_words) {
  extern __shared__ unsigned long long lds[];
  __shared__ int64_t taken[2];
  const double scale = slot[0];
  // markers per exact register sum: the scale keeps every sum of this many below 2^51 quanta
  const int per_pass = slot[2] < 8.0 ? (int)slot[2] : 8;
  // the LDS copy starts zero and each item leaves its words zero again (its flush clears
  // them), so no item clears the copy first
  for (int i = threadIdx.x; i < lds_words; i += blockDim.x) lds[i] = 0ull;
  if (threadIdx.x == 0) taken[0] = (int64_t)atomicAdd(counter, 1u);
  __syncthreads();
  int buf = 0;
  while (true) {
    const int64_t it = taken[buf];
    if (it >= n_items) break;  // every wave of the workgroup leaves together
    // the next item's index, fetched while this one is deposited
    if (threadIdx.x == 0) taken[buf ^ 1] = (int64_t)atomicAdd(counter, 1u);
    const Item item = items[it];
    const mpf_patch p = patches[item.patch];
    int nodes = 1;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) nodes *= (int)p.num_nodes[d];
    const int values = nodes * ncomp;
#define SCATTER(D, K, P)                                                                                        \
  if (ncomp == 1 && (D < 3 || PatchStencil<K, P>::S <= 3))                                                        \
    patch_scatter_runs<D, K, P, T>(p, pos, ps, val, item.first, item.count, (lds_u64 *)lds, scale, per_pass);     \
  else                                                                                                           \
    patch_scatter_item<D, K, P, T>(p, pos, ps, val, vs, ncomp, item.first, item.count, (lds_u64 *)lds, scale)
    PATCH_DISPATCH(NDIM, SCATTER)
#undef SCATTER
    __syncthreads();
    for (int i = threadIdx.x; i < values; i += blockDim.x) {
      const unsigned long long q = lds[i];
      if (q != 0ull) {
        atomicAdd(accum + p.field_offset + i, q);
        lds[i] = 0ull;
      }
    }
    __syncthreads();  // the copy is zero again and the next index is in `taken`
    buf ^= 1;
  }
}

template <class T, int NDIM>
__global__ void __launch_bounds__(256) patches_gather_kernel(const T *__restrict__ pos, int64_t ps, int ncomp,
                                                             const mpf_patch *__restrict__ patches,
                                                             const Item *__restrict__ items, int64_t n_items,
                                                             const double *__restrict__ fields,
                                                             T *__restrict__ res, int64_t rs) {
  // one workgroup per item: the item's markers interpolate its patch's field
  for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const Item item = items[it];
    const mpf_patch p = patches[item.patch];
    const double *field = fields + p.field_offset;
#define GATHER(D, K, P) patch_gather_item<D, K, P, T>(p, pos, ps, ncomp, item.first, item.count, field, res, rs)
    PATCH_DISPATCH(NDIM, GATHER)
#undef GATHER
  }
}

}  // namespace mpf

using namespace mpf;

extern "C" void mpf_patch_plan_destroy(mpf_patch_plan *plan);

extern "C" mpf_patch_plan *mpf_patch_plan_create(int ndim, int64_t n_patches, const mpf_patch *patches,
                                                 int64_t chunk, int64_t ncomp) {
  if (ndim < 1 || ndim > 3 || n_patches < 0 || chunk < 1 || ncomp < 1) {
    set_error("patch plan: need 1 <= ndim <= 3, n_patches >= 0, chunk >= 1, ncomp >= 1");
    return nullptr;
  }
  mpf_patch_plan *plan = new mpf_patch_plan();
  plan->ndim = ndim;
  plan->n_patches = n_patches;
  plan->chunk = chunk;
  std::vector<Item> items;
  int64_t end_values = 0;
  for (int64_t i = 0; i < n_patches; ++i) {
    const mpf_patch &p = patches[i];
    int64_t nodes = 1;
    for (int d = 0; d < ndim; ++d) {
      if (p.num_nodes[d] < 2 || p.code[d] < 0 || p.code[d] > 1) {
        set_error("patch %lld: at least 2 nodes per axis and code 0 or 1", (long long)i);
        delete plan;
        return nullptr;
      }
      nodes *= p.num_nodes[d];
    }
    if (p.kind < LINEAR || p.kind > SPLINE || (p.kind == SPLINE && (p.degree < 0 || p.degree > 3))) {
      set_error("patch %lld: kind 0..2, spline degree 0..3", (long long)i);
      delete plan;
      return nullptr;
    }
    if (nodes > plan->max_nodes) plan->max_nodes = nodes;
    const int64_t top = p.field_offset + nodes * ncomp;
    if (top > end_values) end_values = top;
    plan->markers = p.first_marker + p.num_markers > plan->markers ? p.first_marker + p.num_markers : plan->markers;
    for (int64_t f = 0; f < p.num_markers; f += chunk)
      items.push_back(Item{i, p.first_marker + f, p.num_markers - f < chunk ? p.num_markers - f : chunk});
  }
  plan->field_values = end_values;
  plan->n_items = (int64_t)items.size();
  if (plan->max_nodes * ncomp * 8 > 150 * 1024) {
    set_error("patch plan: a patch of %lld nodes x %lld components exceeds the 150 KB LDS copy",
              (long long)plan->max_nodes, (long long)ncomp);
    delete plan;
    return nullptr;
  }
  if (hipMalloc(&plan->patches, sizeof(mpf_patch) * (n_patches > 0 ? n_patches : 1)) != hipSuccess ||
      hipMalloc(&plan->items, sizeof(Item) * (items.empty() ? 1 : items.size())) != hipSuccess ||
      hipMalloc((void **)&plan->counter, 64) != hipSuccess) {
    set_error("patch plan: device allocation failed");
    delete plan;
    return nullptr;
  }
  if ((n_patches > 0 && hipMemcpy(plan->patches, patches, sizeof(mpf_patch) * n_patches, hipMemcpyHostToDevice) !=
                            hipSuccess) ||
      (!items.empty() &&
       hipMemcpy(plan->items, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("patch plan: upload failed");
    mpf_patch_plan_destroy(plan);
    return nullptr;
  }
  return plan;
}

extern "C" void mpf_patch_plan_destroy(mpf_patch_plan *plan) {
  if (!plan) return;
  (void)hipDeviceSynchronize();
  if (plan->patches) (void)hipFree(plan->patches);
  if (plan->items) (void)hipFree(plan->items);
  if (plan->counter) (void)hipFree(plan->counter);
  delete plan;
}

extern "C" int64_t mpf_patch_plan_items(const mpf_patch_plan *plan) { return plan ? plan->n_items : -1; }

extern "C" int mpf_patches_scatter(const mpf_patch_plan *plan, const void *positions, int64_t pos_stride,
                                   const void *values, int64_t val_stride, int dtype, int64_t ncomp, double *fields,
                                   int block, int workgroups, mpf_context *ctx, mpf_stream stream_) {
  if (!plan) {
    set_error("no patch plan");
    return MPF_ERR_ARGUMENT;
  }
  if (dtype != F32 && dtype != F64) {
    set_error("dtype must be MPF_F32 or MPF_F64");
    return MPF_ERR_ARGUMENT;
  }
  if (block != 256 && block != 512) {
    set_error("patches: block is 256 or 512");
    return MPF_ERR_ARGUMENT;
  }
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  int rc = ensure(ctx->accum, (size_t)(plan->field_values > 0 ? plan->field_values : 1) * 8, true, stream);
  if (rc) return rc;
  rc = prepare_scale(ctx, values, val_stride, dtype, ncomp, plan->markers, 1.0, MPF_SCALE_AUTO, stream);
  if (rc) return rc;
  if (plan->n_items == 0) return MPF_OK;
  hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, stream, plan->counter);
  MPF_LAUNCH_CHECK();
  const size_t lds = (size_t)plan->max_nodes * ncomp * 8;
  if (workgroups < 1) workgroups = 256 * 2;
  {
    ProfileScope prof("patches_scatter", stream);
#define LAUNCH(T, D)                                                                                              \
  hipLaunchKernelGGL((patches_scatter_kernel<T, D>), dim3(workgroups), dim3(block), lds, stream,                  \
                     (const T *)positions, pos_stride, (const T *)values, val_stride, (int)ncomp,                 \
                     (const mpf_patch *)plan->patches, (const Item *)plan->items, plan->n_items, plan->counter, \
                     (unsigned long long *)ctx->accum.ptr, scale_slot(ctx), (int)(lds / 8))
    if (dtype == F32) {
      if (plan->ndim == 1) LAUNCH(float, 1);
      if (plan->ndim == 2) LAUNCH(float, 2);
      if (plan->ndim == 3) LAUNCH(float, 3);
    } else {
      if (plan->ndim == 1) LAUNCH(double, 1);
      if (plan->ndim == 2) LAUNCH(double, 2);
      if (plan->ndim == 3) LAUNCH(double, 3);
    }
#undef LAUNCH
  }
  MPF_LAUNCH_CHECK();
  return finalize_fixed(ctx, fields, plan->field_values, stream);
}

extern "C" int mpf_patches_gather(const mpf_patch_plan *plan, const void *positions, int64_t pos_stride, int dtype,
                                  const double *fields, int64_t ncomp, void *result, int64_t result_stride,
                                  int workgroups, mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  if (!plan) {
    set_error("no patch plan");
    return MPF_ERR_ARGUMENT;
  }
  if (plan->n_items == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  if (workgroups < 1) workgroups = (int)(plan->n_items < 65536 ? plan->n_items : 65536);
  ProfileScope prof("patches_gather", stream);
#define LAUNCH(T, D)                                                                                               \
  hipLaunchKernelGGL((patches_gather_kernel<T, D>), dim3(workgroups), dim3(256), 0, stream, (const T *)positions, \
                     pos_stride, (int)ncomp, (const mpf_patch *)plan->patches, (const Item *)plan->items,         \
                     plan->n_items, fields, (T *)result, result_stride)
  if (dtype == F32) {
    if (plan->ndim == 1) LAUNCH(float, 1);
    if (plan->ndim == 2) LAUNCH(float, 2);
    if (plan->ndim == 3) LAUNCH(float, 3);
  } else {
    if (plan->ndim == 1) LAUNCH(double, 1);
    if (plan->ndim == 2) LAUNCH(double, 2);
    if (plan->ndim == 3) LAUNCH(double, 3);
  }
#undef LAUNCH
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}
