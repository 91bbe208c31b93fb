// mpf_halo.hip -- the on-device seam: halo fill and fold of one block that holds a whole
// periodic axis, and the discrete gradient / divergence of edge fields.
//
// The reference moves halo slabs between MPI ranks (mpiarray update_halos /
// accumulate_halos, called from field.py:601-630); on one GPU the neighbour across a
// periodic seam is the block itself, so fill and fold become copies and adds between the
// two ends of the axis.  Axes are processed one after another, each over the full storage
// of the others (halos included), so a corner value reaches the diagonal in two steps.
#include "mpf_internal.h"

namespace mpf {

// One axis: `outer` = product of the storage extents before the axis, `inner` = product
// after it times ncomp.  One work-item per (outer, inner) column runs over the h layers in
// order, so a short interior (fewer than 2h nodes, where the two seams' layers overlap)
// is folded exactly as layer after layer on one rank.
__global__ void halo_axis_kernel(double *__restrict__ data, int64_t outer, int64_t len, int64_t inner, int64_t h,
                                 int mode, int periodic) {
  const int64_t n = len - 2 * h;
  const int64_t total = outer * inner;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % inner;
    const int64_t o = t / inner;
    double *base = data + o * len * inner + i;
    for (int64_t l = 0; l < h; ++l) {
      double *lo_halo = base + l * inner, *hi_halo = base + (n + h + l) * inner;
      double *lo_in = base + (h + l) * inner, *hi_in = base + (n + l) * inner;
      if (mode == 0) {  // fill
        if (periodic) {
          *lo_halo = *hi_in;
          *hi_halo = *lo_in;
        }
      } else if (periodic) {  // fold: add across the seam
        *lo_in += *hi_halo;
        *hi_in += *lo_halo;
      }
    }
    if (mode == 1)  // then clear the halos
      for (int64_t l = 0; l < h; ++l) {
        base[l * inner] = 0.0;
        base[(n + h + l) * inner] = 0.0;
      }
  }
}

static int halo(int mode, int ndim, const int64_t *shape, int64_t ncomp, int64_t h, const int64_t *periodic,
                double *data, hipStream_t stream) {
  if (ndim < 1 || ndim > 3 || ncomp < 1 || h < 0) {
    set_error("halo: need 1 <= ndim <= 3, ncomp >= 1, halo >= 0");
    return MPF_ERR_ARGUMENT;
  }
  if (h == 0) return MPF_OK;
  for (int d = 0; d < ndim; ++d) {
    if (shape[d] - 2 * h < h) {
      set_error("halo: axis %d stores %lld nodes, fewer than halo interior nodes plus its halos", d,
                (long long)shape[d]);
      return MPF_ERR_ARGUMENT;
    }
  }
  for (int axis = 0; axis < ndim; ++axis) {
    int64_t outer = 1, inner = ncomp;
    for (int d = 0; d < axis; ++d) outer *= shape[d];
    for (int d = axis + 1; d < ndim; ++d) inner *= shape[d];
    const int64_t total = outer * inner;
    hipLaunchKernelGGL(halo_axis_kernel, dim3(blocks_for(total, 256, 1 << 16)), dim3(256), 0, stream, data, outer,
                       shape[axis], inner, h, mode, (int)(periodic[axis] != 0));
    MPF_LAUNCH_CHECK();
  }
  return MPF_OK;
}

struct Shape3 {
  int64_t n[3];
  int64_t stride[3];
  int periodic[3];
  int ndim;
  int64_t total;
};

static Shape3 shape3(int ndim, const int64_t *shape, const int64_t *periodic) {
  Shape3 s;
  s.ndim = ndim;
  s.total = 1;
  for (int d = 0; d < 3; ++d) {
    s.n[d] = d < ndim ? shape[d] : 1;
    s.periodic[d] = d < ndim ? (int)(periodic[d] != 0) : 0;
  }
  int64_t st = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    s.stride[d] = st;
    st *= s.n[d];
  }
  for (int d = ndim; d < 3; ++d) s.stride[d] = 0;
  s.total = st;
  return s;
}

// G^T j (edges.py:240-259), gathered per node: out = sum over axes of j(prev edge) - j(own edge).
__global__ void divergence_kernel(const double *__restrict__ edges, double *__restrict__ out, Shape3 s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < s.total; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int axis = 0; axis < s.ndim; ++axis) {
      const int64_t k = (i / s.stride[axis]) % s.n[axis];
      const double own = edges[i * s.ndim + axis];
      if (s.periodic[axis]) {
        const int64_t prev = k == 0 ? i + (s.n[axis] - 1) * s.stride[axis] : i - s.stride[axis];
        acc += edges[prev * s.ndim + axis] - own;
      } else {
        if (k < s.n[axis] - 1) acc -= own;
        if (k > 0) acc += edges[(i - s.stride[axis]) * s.ndim + axis];
      }
    }
    out[i] = acc;
  }
}

// G phi (edges.py:218-237).
__global__ void gradient_kernel(const double *__restrict__ nodes, double *__restrict__ out, Shape3 s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < s.total; i += (int64_t)gridDim.x * blockDim.x) {
    for (int axis = 0; axis < s.ndim; ++axis) {
      const int64_t k = (i / s.stride[axis]) % s.n[axis];
      double v = 0.0;
      if (s.periodic[axis]) {
        const int64_t next = k == s.n[axis] - 1 ? i - (s.n[axis] - 1) * s.stride[axis] : i + s.stride[axis];
        v = nodes[next] - nodes[i];
      } else if (k < s.n[axis] - 1) {
        v = nodes[i + s.stride[axis]] - nodes[i];
      }
      out[i * s.ndim + axis] = v;
    }
  }
}

}  // namespace mpf

using namespace mpf;

extern "C" int mpf_halo_fill(int ndim, const int64_t *storage_shape, int64_t ncomp, int64_t halo_width,
                             const int64_t *periodic, double *data, mpf_stream stream) {
  return halo(0, ndim, storage_shape, ncomp, halo_width, periodic, data, (hipStream_t)stream);
}

extern "C" int mpf_halo_fold(int ndim, const int64_t *storage_shape, int64_t ncomp, int64_t halo_width,
                             const int64_t *periodic, double *data, mpf_stream stream) {
  return halo(1, ndim, storage_shape, ncomp, halo_width, periodic, data, (hipStream_t)stream);
}

extern "C" int mpf_divergence(int ndim, const int64_t *shape, const int64_t *periodic, const double *edges,
                              double *out, mpf_stream stream) {
  if (ndim < 1 || ndim > 3) {
    set_error("divergence: ndim must be 1, 2 or 3");
    return MPF_ERR_ARGUMENT;
  }
  Shape3 s = shape3(ndim, shape, periodic);
  hipLaunchKernelGGL(divergence_kernel, dim3(blocks_for(s.total, 256, 1 << 16)), dim3(256), 0, (hipStream_t)stream,
                     edges, out, s);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

extern "C" int mpf_gradient(int ndim, const int64_t *shape, const int64_t *periodic, const double *nodes, double *out,
                            mpf_stream stream) {
  if (ndim < 1 || ndim > 3) {
    set_error("gradient: ndim must be 1, 2 or 3");
    return MPF_ERR_ARGUMENT;
  }
  Shape3 s = shape3(ndim, shape, periodic);
  hipLaunchKernelGGL(gradient_kernel, dim3(blocks_for(s.total, 256, 1 << 16)), dim3(256), 0, (hipStream_t)stream,
                     nodes, out, s);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}
