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

// One patch's per-axis stencils (kind/degree of the patch), as the direct kernels compute them.
__device__ inline int patch_axis(const mpf_patch &p, int d, double x, int64_t *node, double *w) {
  const double rel = (x - p.lower[d]) * (1.0 / p.cell_size[d]);
  if (p.kind == LINEAR) return linear_axis(rel, p.num_nodes[d], (int)p.code[d], node, w);
  if (p.kind == QUADRATIC) return quadratic_axis(rel, p.num_nodes[d], (int)p.code[d], node, w);
  return spline_axis(rel, p.num_nodes[d], (int)p.code[d], p.degree, node, w);
}

__global__ void zero_counter(unsigned int *c) { *c = 0u; }

template <class T>
__global__ void __launch_bounds__(512) patches_scatter_kernel(const T *__restrict__ pos, int64_t ps,
                                                              const T *__restrict__ val, int64_t vs, int ncomp,
                                                              int ndim, const mpf_patch *__restrict__ patches,
                                                              const Item *__restrict__ items, int64_t n_items,
                                                              unsigned int *__restrict__ counter,
                                                              unsigned long long *__restrict__ accum,
                                                              const double *__restrict__ slot) {
  extern __shared__ unsigned long long lds[];
  __shared__ int64_t current;
  const double scale = slot[0];
  while (true) {
    if (threadIdx.x == 0) current = (int64_t)atomicAdd(counter, 1u);
    __syncthreads();
    const int64_t it = current;
    if (it >= n_items) break;  // every wave of the workgroup leaves together
    const Item item = items[it];
    const mpf_patch p = patches[item.patch];
    int64_t nodes = 1;
    for (int d = 0; d < ndim; ++d) nodes *= p.num_nodes[d];
    const int64_t values = nodes * ncomp;
    for (int64_t i = threadIdx.x; i < values; i += blockDim.x) lds[i] = 0ull;
    __syncthreads();
    for (int64_t m = item.first + threadIdx.x; m < item.first + item.count; m += blockDim.x) {
      bool live = false;
      for (int c = 0; c < ncomp; ++c) live |= (double)val[c * vs + m] != 0.0;
      if (!live) continue;
      int64_t node[3][4];
      double w[3][4];
      int size[3] = {1, 1, 1};
      node[1][0] = node[2][0] = 0;
      w[1][0] = w[2][0] = 1.0;
      for (int d = 0; d < ndim; ++d) size[d] = patch_axis(p, d, (double)pos[d * ps + m], node[d], w[d]);
      const int64_t s1 = ndim == 3 ? p.num_nodes[2] : 1;
      const int64_t s0 = ndim == 3 ? p.num_nodes[1] * p.num_nodes[2] : (ndim == 2 ? p.num_nodes[1] : 1);
      for (int a = 0; a < size[0]; ++a)
        for (int b = 0; b < size[1]; ++b) {
          const double wab = ndim >= 2 ? w[0][a] * w[1][b] : w[0][a];
          for (int e = 0; e < size[2]; ++e) {
            const double weight = ndim == 3 ? wab * w[2][e] : wab;
            const int64_t at =
                (node[0][a] * s0 + (ndim >= 2 ? node[1][b] * s1 : 0) + (ndim == 3 ? node[2][e] : 0)) * ncomp;
            for (int c = 0; c < ncomp; ++c)
              atomicAdd(&lds[at + c], to_fixed(weight * (double)val[c * vs + m] * scale));
          }
        }
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < values; i += blockDim.x) {
      const unsigned long long q = lds[i];
      if (q != 0ull) atomicAdd(accum + p.field_offset + i, q);
    }
    __syncthreads();  // the LDS copy and `current` are reused by the next item
  }
}

template <class T>
__global__ void __launch_bounds__(256) patches_gather_kernel(const T *__restrict__ pos, int64_t ps, int ncomp,
                                                             int ndim, const mpf_patch *__restrict__ patches,
                                                             const Item *__restrict__ items, int64_t n_items,
                                                             const double *__restrict__ fields,
                                                             T *__restrict__ res, int64_t rs) {
  // one workgroup per item: the item's markers interpolate its patch's field
  for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const Item item = items[it];
    const mpf_patch p = patches[item.patch];
    const double *field = fields + p.field_offset;
    const int64_t s1 = ndim == 3 ? p.num_nodes[2] : 1;
    const int64_t s0 = ndim == 3 ? p.num_nodes[1] * p.num_nodes[2] : (ndim == 2 ? p.num_nodes[1] : 1);
    for (int64_t m = item.first + threadIdx.x; m < item.first + item.count; m += blockDim.x) {
      int64_t node[3][4];
      double w[3][4];
      int size[3] = {1, 1, 1};
      node[1][0] = node[2][0] = 0;
      w[1][0] = w[2][0] = 1.0;
      for (int d = 0; d < ndim; ++d) size[d] = patch_axis(p, d, (double)pos[d * ps + m], node[d], w[d]);
      for (int c = 0; c < ncomp; ++c) {
        double total = 0.0;
        for (int a = 0; a < size[0]; ++a)
          for (int b = 0; b < size[1]; ++b) {
            const double wab = ndim >= 2 ? w[0][a] * w[1][b] : w[0][a];
            for (int e = 0; e < size[2]; ++e) {
              const double weight = ndim == 3 ? wab * w[2][e] : wab;
              total += weight * field[(node[0][a] * s0 + (ndim >= 2 ? node[1][b] * s1 : 0) +
                                       (ndim == 3 ? node[2][e] : 0)) * ncomp + c];
            }
          }
        res[c * rs + m] = (T)total;
      }
    }
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
  if (!ctx) ctx = default_context();
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
    if (dtype == F32)
      hipLaunchKernelGGL(patches_scatter_kernel<float>, dim3(workgroups), dim3(block), lds, stream,
                         (const float *)positions, pos_stride, (const float *)values, val_stride, (int)ncomp,
                         plan->ndim, (const mpf_patch *)plan->patches, (const Item *)plan->items, plan->n_items,
                         plan->counter, (unsigned long long *)ctx->accum.ptr, scale_slot(ctx));
    else
      hipLaunchKernelGGL(patches_scatter_kernel<double>, dim3(workgroups), dim3(block), lds, stream,
                         (const double *)positions, pos_stride, (const double *)values, val_stride, (int)ncomp,
                         plan->ndim, (const mpf_patch *)plan->patches, (const Item *)plan->items, plan->n_items,
                         plan->counter, (unsigned long long *)ctx->accum.ptr, scale_slot(ctx));
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
  if (dtype == F32)
    hipLaunchKernelGGL(patches_gather_kernel<float>, dim3(workgroups), dim3(256), 0, stream, (const float *)positions,
                       pos_stride, (int)ncomp, plan->ndim, (const mpf_patch *)plan->patches,
                       (const Item *)plan->items, plan->n_items, fields, (float *)result, result_stride);
  else
    hipLaunchKernelGGL(patches_gather_kernel<double>, dim3(workgroups), dim3(256), 0, stream,
                       (const double *)positions, pos_stride, (int)ncomp, plan->ndim,
                       (const mpf_patch *)plan->patches, (const Item *)plan->items, plan->n_items, fields,
                       (double *)result, result_stride);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}
