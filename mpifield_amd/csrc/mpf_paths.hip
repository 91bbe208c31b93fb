// mpf_paths.hip -- the charge-conserving path-current deposits, one work-item per path.
//
// WHITNEY (deposit_paths_{1,2,3}d_pyccel.py): the path, clamped to the grid on bounded
// axes, is cut at every grid plane it crosses; each piece adds q * its length along an
// axis to the edges of its cell along that axis, weighted by the piece's midpoint across
// the cell (2D) or its average over the piece, midpoint +- du dw / 12 (3D,
// whitney_cell_pyccel.py:31-49).
// QUADRATIC_EDGES / CUBIC_EDGES (deposit_paths_spline_{1,2,3}d_pyccel.py): stretches of at
// most 4 cells, cut at the knot planes and walls (spline_edges_pyccel.py:180-230), each
// piece integrated exactly by Gauss-Legendre quadrature of (p-1+(ndim-1)p)//2+1 points.
// The segment split runs in float64 on the device, in the reference's operation order, so
// that the pieces are the reference's; the edge sums go through float64 atomics
// (MODE_ATOMIC) or int64 fixed point (MODE_FIXED, bit-reproducible).
#include "mpf_internal.h"

namespace mpf {

__device__ __forceinline__ double clamp_to_axis(double rel, int64_t num, int code) {
  if (code == 1) return rel;
  return rel < 0.0 ? 0.0 : (rel > (double)num - 1.0 ? (double)num - 1.0 : rel);
}

__device__ __forceinline__ double plane_time(double plane, double rel, double delta) {
  return delta != 0.0 ? (plane - rel) / delta : 2.0;
}

__device__ __forceinline__ double edge_path_weight(double fu, double fw, double du, double dw, int uu, int uw) {
  return (uu == 1 ? fu : 1.0 - fu) * (uw == 1 ? fw : 1.0 - fw) + (uu == uw ? du * dw / 12.0 : -du * dw / 12.0);
}

template <int MODE>
__device__ __forceinline__ void deposit(double *edges, unsigned long long *accum, int64_t at, double v, double scale) {
  if (MODE == MODE_ATOMIC)
    unsafeAtomicAdd(edges + at, v);
  else
    atomicAdd(accum + at, to_fixed(v * scale));
}

template <class T, int NDIM, int MODE>
__global__ void __launch_bounds__(256) whitney_kernel(const T *__restrict__ st, int64_t ss, const T *__restrict__ en,
                                                      int64_t es, const T *__restrict__ q, int64_t n, Grid g,
                                                      double *__restrict__ edges,
                                                      unsigned long long *__restrict__ accum,
                                                      const double *__restrict__ slot) {
  const double scale = MODE == MODE_FIXED ? slot[0] : 1.0;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const double charge = (double)q[p];
    if (charge == 0.0) continue;
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    bool finite = true;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      a[d] = ((double)st[d * ss + p] - g.lower[d]) * g.inverse[d];
      b[d] = ((double)en[d * es + p] - g.lower[d]) * g.inverse[d];
      finite = finite && a[d] > -1.0e300 && a[d] < 1.0e300 && b[d] > -1.0e300 && b[d] < 1.0e300;
    }
    if (!finite) continue;
    double delta[3] = {0, 0, 0}, step[3] = {0, 0, 0}, plane[3] = {0, 0, 0}, x0[3] = {0, 0, 0}, x1[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      delta[d] = b[d] - a[d];
      step[d] = 1.0;
      plane[d] = floor(a[d]) + 1.0;
      if (delta[d] < 0.0) {
        step[d] = -1.0;
        plane[d] = ceil(a[d]) - 1.0;
      }
      x0[d] = clamp_to_axis(a[d], g.num[d], g.code[d]);
    }
    double time = 0.0;
    while (time < 1.0) {
      double following = 1.0;
      double times[3] = {2.0, 2.0, 2.0};
      if (NDIM == 1) {
        following = plane_time(plane[0], a[0], delta[0]);
        if (following < 1.0) {
          x1[0] = clamp_to_axis(a[0] + following * delta[0], g.num[0], g.code[0]);
          plane[0] += step[0];
        } else {
          following = 1.0;
          x1[0] = clamp_to_axis(b[0], g.num[0], g.code[0]);
        }
      } else {
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          times[d] = plane_time(plane[d], a[d], delta[d]);
          if (times[d] < following) following = times[d];
        }
#pragma unroll
        for (int d = 0; d < NDIM; ++d)
          x1[d] = following < 1.0 ? clamp_to_axis(a[d] + following * delta[d], g.num[d], g.code[d])
                                  : clamp_to_axis(b[d], g.num[d], g.code[d]);
#pragma unroll
        for (int d = 0; d < NDIM; ++d)
          if (times[d] <= following) plane[d] += step[d];
      }
      int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      double f[3] = {0, 0, 0}, dd[3] = {0, 0, 0}, along[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        double mid = 0.5 * (x0[d] + x1[d]);
        int64_t nd[2];
        double wd[2];
        linear_axis(mid, g.num[d], g.code[d], nd, wd);  // periodic_position + left/right node + fraction
        lo[d] = nd[0];
        hi[d] = nd[1];
        f[d] = wd[1];
        dd[d] = x1[d] - x0[d];
        along[d] = charge * dd[d];
      }
      if (NDIM == 1) {
        deposit<MODE>(edges, accum, lo[0], charge * (x1[0] - x0[0]), scale);
      } else if (NDIM == 2) {
        const int64_t s = g.stride[0];
        deposit<MODE>(edges, accum, (lo[0] * s + lo[1]) * 2 + 0, (1.0 - f[1]) * along[0], scale);
        deposit<MODE>(edges, accum, (lo[0] * s + hi[1]) * 2 + 0, f[1] * along[0], scale);
        deposit<MODE>(edges, accum, (lo[0] * s + lo[1]) * 2 + 1, (1.0 - f[0]) * along[1], scale);
        deposit<MODE>(edges, accum, (hi[0] * s + lo[1]) * 2 + 1, f[0] * along[1], scale);
      } else {
        const int64_t s0 = g.stride[0], s1 = g.stride[1];
#define AT(i, j, k, c) (((i)*s0 + (j)*s1 + (k)) * 3 + (c))
        deposit<MODE>(edges, accum, AT(lo[0], lo[1], lo[2], 0), along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 0, 0), scale);
        deposit<MODE>(edges, accum, AT(lo[0], hi[1], lo[2], 0), along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 1, 0), scale);
        deposit<MODE>(edges, accum, AT(lo[0], lo[1], hi[2], 0), along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 0, 1), scale);
        deposit<MODE>(edges, accum, AT(lo[0], hi[1], hi[2], 0), along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 1, 1), scale);
        deposit<MODE>(edges, accum, AT(lo[0], lo[1], lo[2], 1), along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 0, 0), scale);
        deposit<MODE>(edges, accum, AT(hi[0], lo[1], lo[2], 1), along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 1, 0), scale);
        deposit<MODE>(edges, accum, AT(lo[0], lo[1], hi[2], 1), along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 0, 1), scale);
        deposit<MODE>(edges, accum, AT(hi[0], lo[1], hi[2], 1), along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 1, 1), scale);
        deposit<MODE>(edges, accum, AT(lo[0], lo[1], lo[2], 2), along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 0, 0), scale);
        deposit<MODE>(edges, accum, AT(hi[0], lo[1], lo[2], 2), along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 1, 0), scale);
        deposit<MODE>(edges, accum, AT(lo[0], hi[1], lo[2], 2), along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 0, 1), scale);
        deposit<MODE>(edges, accum, AT(hi[0], hi[1], lo[2], 2), along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 1, 1), scale);
#undef AT
      }
      time = following;
#pragma unroll
      for (int d = 0; d < NDIM; ++d) x0[d] = x1[d];
    }
  }
}

// ---- spline edges ---------------------------------------------------------------------

__device__ __forceinline__ int64_t floor_int(double x) {
  int64_t whole = (int64_t)x;
  if (x < (double)whole) whole -= 1;
  return whole;
}

// spline_edges_pyccel.py:86-138: the `degree` edges along an axis around rel, weights
// B_{p-1}(rel - k - 1/2); on a bounded axis an edge past a wall does not exist.
__device__ inline void along_stencil(double rel, int64_t num, int code, int degree, int64_t *edge, double *w,
                                     int *exists) {
  if (code == 2) {
    clamped_cell(rel, num, degree, 1, edge, w);
    for (int a = 0; a < degree; ++a) exists[a] = 1;
    return;
  }
  const int lower = degree - 1;
  double r = rel;
  if (code == 1 && !(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
  const double shifted = r - 0.5;
  int64_t middle;
  if (lower % 2 == 0) {
    middle = floor_int(shifted + 0.5);
    if (code == 0 && middle > num - 2) middle = num - 2;
  } else {
    middle = floor_int(shifted);
  }
  const double o = shifted - (double)middle;
  if (lower == 0) {
    w[0] = 1.0;
  } else if (lower == 1) {
    w[0] = 1.0 - o;
    w[1] = o;
  } else {
    w[0] = 0.5 * (0.5 - o) * (0.5 - o);
    w[1] = 0.75 - o * o;
    w[2] = 0.5 * (0.5 + o) * (0.5 + o);
  }
  const int64_t first = middle - lower / 2;
  for (int a = 0; a < degree; ++a) {
    const int64_t k = first + a;
    if (code == 1) {
      edge[a] = (k + 2 * num) % num;
      exists[a] = 1;
    } else if (k < 0 || k > num - 2) {
      edge[a] = 0;
      exists[a] = 0;
    } else {
      edge[a] = k;
      exists[a] = 1;
    }
  }
}

__device__ inline void gauss_rule(int count, double *point, double *weight) {
  if (count == 1) {
    point[0] = 0.5;
    weight[0] = 1.0;
  } else if (count == 2) {
    point[0] = 0.21132486540518713;
    point[1] = 0.7886751345948129;
    weight[0] = 0.5;
    weight[1] = 0.5;
  } else if (count == 3) {
    point[0] = 0.1127016653792583;
    point[1] = 0.5;
    point[2] = 0.8872983346207417;
    weight[0] = 0.2777777777777778;
    weight[1] = 0.4444444444444444;
    weight[2] = 0.2777777777777778;
  } else if (count == 4) {
    point[0] = 0.06943184420297371;
    point[1] = 0.33000947820757187;
    point[2] = 0.6699905217924281;
    point[3] = 0.9305681557970262;
    weight[0] = 0.17392742256872692;
    weight[1] = 0.3260725774312731;
    weight[2] = 0.3260725774312731;
    weight[3] = 0.17392742256872692;
  } else {
    point[0] = 0.04691007703066802;
    point[1] = 0.23076534494715845;
    point[2] = 0.5;
    point[3] = 0.7692346550528415;
    point[4] = 0.9530899229693319;
    weight[0] = 0.11846344252809454;
    weight[1] = 0.2393143352496832;
    weight[2] = 0.28444444444444444;
    weight[3] = 0.2393143352496832;
    weight[4] = 0.11846344252809454;
  }
}

// spline_edges_pyccel.py:180-230 path_cuts, with the reference's insertion sort.
template <int NDIM>
__device__ inline int path_cuts(const double *r0, const double *r1, double shift, const Grid &g, double *cuts) {
  int count = 1;
  cuts[0] = 0.0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const double delta = r1[d] - r0[d];
    if (delta == 0.0) continue;
    const double low = delta > 0.0 ? r0[d] : r1[d];
    const double high = delta > 0.0 ? r1[d] : r0[d];
    const double offset = g.code[d] == 2 ? 0.0 : shift;
    double plane = (double)floor_int(low - offset) + 1.0 + offset;
    while (plane < high) {
      const double t = (plane - r0[d]) / delta;
      if (t > 0.0 && t < 1.0 && count < 31) cuts[count++] = t;
      plane += 1.0;
    }
    if (g.code[d] != 1) {
      double t = (0.0 - r0[d]) / delta;
      if (t > 0.0 && t < 1.0 && count < 31) cuts[count++] = t;
      t = ((double)g.num[d] - 1.0 - r0[d]) / delta;
      if (t > 0.0 && t < 1.0 && count < 31) cuts[count++] = t;
    }
  }
  cuts[count++] = 1.0;
  for (int i = 2; i < count - 1; ++i) {
    const double value = cuts[i];
    int j = i - 1;
    while (j > 0 && cuts[j] > value) {
      cuts[j + 1] = cuts[j];
      --j;
    }
    cuts[j + 1] = value;
  }
  return count;
}

template <class T, int NDIM, int MODE>
__global__ void __launch_bounds__(128) spline_paths_kernel(const T *__restrict__ st, int64_t ss,
                                                           const T *__restrict__ en, int64_t es,
                                                           const T *__restrict__ q, int64_t n, Grid g, int degree,
                                                           double *__restrict__ edges,
                                                           unsigned long long *__restrict__ accum,
                                                           const double *__restrict__ slot) {
  const double scale = MODE == MODE_FIXED ? slot[0] : 1.0;
  const double shift = degree % 2 == 0 ? 0.5 : 0.0;
  const int num_points = (degree - 1 + (NDIM - 1) * degree) / 2 + 1;
  double point[5], gw[5];
  gauss_rule(num_points, point, gw);
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const double charge = (double)q[p];
    if (charge == 0.0) continue;
    bool finite = true;
    double r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0}, extent = 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const double s = (double)st[d * ss + p], e = (double)en[d * es + p];
      finite = finite && (s - s == 0.0) && (e - e == 0.0);
      r0[d] = (s - g.lower[d]) * g.inverse[d];
      r1[d] = (e - g.lower[d]) * g.inverse[d];
      const double span = r1[d] > r0[d] ? r1[d] - r0[d] : r0[d] - r1[d];
      if (span > extent) extent = span;
    }
    if (!finite) continue;
    const int64_t stretches = (int64_t)(extent / 4.0) + 1;
    for (int64_t s = 0; s < stretches; ++s) {
      double a0[3] = {0, 0, 0}, a1[3] = {0, 0, 0}, cuts[32];
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        a0[d] = r0[d] + (r1[d] - r0[d]) * (double)s / (double)stretches;
        a1[d] = r0[d] + (r1[d] - r0[d]) * (double)(s + 1) / (double)stretches;
      }
      const int num_cuts = path_cuts<NDIM>(a0, a1, shift, g, cuts);
      for (int piece = 0; piece < num_cuts - 1; ++piece) {
        const double t0 = cuts[piece], t1 = cuts[piece + 1];
        if (t1 <= t0) continue;
        double p0[3] = {0, 0, 0}, stp[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          double begin = a0[d] + t0 * (a1[d] - a0[d]);
          double finish = a0[d] + t1 * (a1[d] - a0[d]);
          if (g.code[d] != 1) {
            const double top = (double)(g.num[d] - 1);
            if (begin < 0.0) begin = 0.0;
            if (begin > top) begin = (double)g.num[d] - 1.0;
            if (finish < 0.0) finish = 0.0;
            if (finish > top) finish = (double)g.num[d] - 1.0;
          }
          p0[d] = begin;
          stp[d] = finish - begin;
        }
        for (int gp = 0; gp < num_points; ++gp) {
          double rel[3];
#pragma unroll
          for (int d = 0; d < NDIM; ++d) rel[d] = p0[d] + point[gp] * stp[d];
#pragma unroll
          for (int axis = 0; axis < NDIM; ++axis) {
            const double amount = charge * gw[gp] * stp[axis];
            if (amount == 0.0) continue;
            int64_t idx[3][4];
            double w[3][4];
            int ex[4] = {1, 1, 1, 1}, size[3] = {1, 1, 1};
            idx[1][0] = idx[2][0] = 0;
            w[1][0] = w[2][0] = 1.0;
#pragma unroll
            for (int d = 0; d < NDIM; ++d) {
              if (d == axis) {
                along_stencil(rel[d], g.num[d], g.code[d], degree, idx[d], w[d], ex);
                size[d] = degree;
              } else {
                spline_axis(rel[d], g.num[d], g.code[d], degree, idx[d], w[d]);
                size[d] = degree + 1;
              }
            }
            for (int a = 0; a < size[0]; ++a) {
              if (axis == 0 && !ex[a]) continue;
              for (int b = 0; b < size[1]; ++b) {
                if (axis == 1 && !ex[b]) continue;
                for (int e = 0; e < size[2]; ++e) {
                  if (axis == 2 && !ex[e]) continue;
                  double v = amount * w[0][a];
                  if (NDIM >= 2) v = v * w[1][b];
                  if (NDIM == 3) v = v * w[2][e];
                  const int64_t at =
                      (idx[0][a] * g.stride[0] + idx[1][b] * g.stride[1] + idx[2][e] * g.stride[2]) * NDIM + axis;
                  deposit<MODE>(edges, accum, at, v, scale);
                }
              }
            }
          }
        }
      }
    }
  }
}

// The fixed-point bound of a path deposit: every edge value is at most
// sum_paths |q| * (length in cells + 1).
template <class T>
__global__ void path_bound_kernel(const T *__restrict__ st, int64_t ss, const T *__restrict__ en, int64_t es,
                                  const T *__restrict__ q, int64_t n, Grid g, unsigned long long *__restrict__ out) {
  double m = 0.0;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    double len = 0.0;
    for (int d = 0; d < g.ndim; ++d) len += fabs(((double)en[d * es + p] - (double)st[d * ss + p]) * g.inverse[d]);
    const double b = fabs((double)q[p]) * (len + 2.0);
    if (b > m && b < 1.0e300) m = b;
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

template <class T, int NDIM>
static void launch_paths(int kind, int mode, const void *st, int64_t ss, const void *en, int64_t es, const void *q,
                         int64_t n, const Grid &g, int degree, double *edges, unsigned long long *accum,
                         const double *slot, hipStream_t stream) {
  if (kind == 0) {
    const int threads = 256, blocks = blocks_for(n, threads, 1 << 16);
    if (mode == MODE_ATOMIC)
      hipLaunchKernelGGL((whitney_kernel<T, NDIM, MODE_ATOMIC>), dim3(blocks), dim3(threads), 0, stream,
                         (const T *)st, ss, (const T *)en, es, (const T *)q, n, g, edges, accum, slot);
    else
      hipLaunchKernelGGL((whitney_kernel<T, NDIM, MODE_FIXED>), dim3(blocks), dim3(threads), 0, stream,
                         (const T *)st, ss, (const T *)en, es, (const T *)q, n, g, edges, accum, slot);
  } else {
    const int threads = 128, blocks = blocks_for(n, threads, 1 << 16);
    if (mode == MODE_ATOMIC)
      hipLaunchKernelGGL((spline_paths_kernel<T, NDIM, MODE_ATOMIC>), dim3(blocks), dim3(threads), 0, stream,
                         (const T *)st, ss, (const T *)en, es, (const T *)q, n, g, degree, edges, accum, slot);
    else
      hipLaunchKernelGGL((spline_paths_kernel<T, NDIM, MODE_FIXED>), dim3(blocks), dim3(threads), 0, stream,
                         (const T *)st, ss, (const T *)en, es, (const T *)q, n, g, degree, edges, accum, slot);
  }
}

template <class T>
static void dispatch_paths(int kind, int ndim, int mode, const void *st, int64_t ss, const void *en, int64_t es,
                           const void *q, int64_t n, const Grid &g, int degree, double *edges,
                           unsigned long long *accum, const double *slot, hipStream_t stream) {
  if (ndim == 1) launch_paths<T, 1>(kind, mode, st, ss, en, es, q, n, g, degree, edges, accum, slot, stream);
  if (ndim == 2) launch_paths<T, 2>(kind, mode, st, ss, en, es, q, n, g, degree, edges, accum, slot, stream);
  if (ndim == 3) launch_paths<T, 3>(kind, mode, st, ss, en, es, q, n, g, degree, edges, accum, slot, stream);
}

}  // namespace mpf

using namespace mpf;

extern "C" int mpf_deposit_paths(int kind, int degree, int ndim, const void *start, int64_t start_stride,
                                 const void *end, int64_t end_stride, const void *charges, int dtype, int64_t n,
                                 double *edges, const double *lower, const double *cell_size,
                                 const int64_t *num_nodes, const int64_t *code, int mode, int scale_exp,
                                 mpf_context *ctx, mpf_stream stream_) {
  if (kind != 0 && kind != 1) {
    set_error("path kind must be 0 (Whitney) or 1 (spline edges), got %d", kind);
    return MPF_ERR_ARGUMENT;
  }
  if (kind == 1 && (degree < 1 || degree > 3)) {
    set_error("spline edge degrees are 1 to 3, got %d", degree);
    return MPF_ERR_ARGUMENT;
  }
  if (dtype != F32 && dtype != F64) {
    set_error("dtype must be MPF_F32 or MPF_F64, got %d", dtype);
    return MPF_ERR_ARGUMENT;
  }
  if (mode != MODE_ATOMIC && mode != MODE_FIXED) {
    set_error("mode must be MPF_MODE_ATOMIC or MPF_MODE_FIXED, got %d", mode);
    return MPF_ERR_ARGUMENT;
  }
  if (n < 0) {
    set_error("negative path count");
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  const int deg = kind == 0 ? 1 : degree;
  int rc = make_grid(g, ndim, kind == 0 ? LINEAR : SPLINE, deg, lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context();
  if (!ctx) return MPF_ERR_HIP;
  const int64_t count = g.extent[0] * (ndim > 1 ? g.extent[1] : 1) * (ndim > 2 ? g.extent[2] : 1) * ndim;
  unsigned long long *accum = nullptr;
  const double *slot = nullptr;
  if (mode == MODE_FIXED) {
    rc = ensure(ctx->accum, (size_t)count * 8, true, stream);
    if (rc) return rc;
    rc = ensure(ctx->scalars, 64, true, stream);
    if (rc) return rc;
    unsigned long long *maxbits = (unsigned long long *)ctx->scalars.ptr;
    if (scale_exp == MPF_SCALE_AUTO) {
      rc = zero_async(maxbits, 8, stream);
      if (rc) return rc;
      const int threads = 256, blocks = blocks_for(n, threads, 4096);
      if (dtype == F32)
        hipLaunchKernelGGL(path_bound_kernel<float>, dim3(blocks), dim3(threads), 0, stream, (const float *)start,
                           start_stride, (const float *)end, end_stride, (const float *)charges, n, g, maxbits);
      else
        hipLaunchKernelGGL(path_bound_kernel<double>, dim3(blocks), dim3(threads), 0, stream, (const double *)start,
                           start_stride, (const double *)end, end_stride, (const double *)charges, n, g, maxbits);
      MPF_LAUNCH_CHECK();
    }
    // scale from n * max(|q| (length + 2)) read back from the slot the bound kernel filled
    rc = prepare_scale(ctx, nullptr, 0, dtype, 1, n, 1.0, scale_exp, stream);
    if (rc) return rc;
    accum = (unsigned long long *)ctx->accum.ptr;
    slot = scale_slot(ctx);
  }
  if (dtype == F32)
    dispatch_paths<float>(kind, ndim, mode, start, start_stride, end, end_stride, charges, n, g, deg, edges, accum,
                          slot, stream);
  else
    dispatch_paths<double>(kind, ndim, mode, start, start_stride, end, end_stride, charges, n, g, deg, edges, accum,
                           slot, stream);
  MPF_LAUNCH_CHECK();
  if (mode == MODE_FIXED) return finalize_fixed(ctx, edges, count, stream);
  return MPF_OK;
}

#define GEOM1 MPF_GEOM_1D
#define GEOM2 MPF_GEOM_2D
#define GEOM3 MPF_GEOM_3D
#define P1 const double lo[1] = {lower_x}; const double cs[1] = {cell_size_x}; const int64_t nn[1] = {num_x}; \
  const int64_t pc[1] = {periodic_x};
#define P2 const double lo[2] = {lower_x, lower_y}; const double cs[2] = {cell_size_x, cell_size_y}; \
  const int64_t nn[2] = {num_x, num_y}; const int64_t pc[2] = {periodic_x, periodic_y};
#define P3 const double lo[3] = {lower_x, lower_y, lower_z}; \
  const double cs[3] = {cell_size_x, cell_size_y, cell_size_z}; const int64_t nn[3] = {num_x, num_y, num_z}; \
  const int64_t pc[3] = {periodic_x, periodic_y, periodic_z};

#define PATH_ENTRY(D)                                                                                               \
  extern "C" int mpf_deposit_paths_##D##d(MPF_PATHS_IN, GEOM##D, int mode, MPF_TAIL) {                     \
    P##D return mpf_deposit_paths(0, 1, D, start, start_stride, end, end_stride, charges, dtype, n, edges, lo, cs,  \
                                  nn, pc, mode, MPF_SCALE_AUTO, ctx, stream);                                      \
  }                                                                                                                \
  extern "C" int mpf_deposit_paths_spline_##D##d(MPF_PATHS_IN, GEOM##D, int degree, int mode, MPF_TAIL) {  \
    P##D return mpf_deposit_paths(1, degree, D, start, start_stride, end, end_stride, charges, dtype, n, edges, lo, \
                                  cs, nn, pc, mode, MPF_SCALE_AUTO, ctx, stream);                                  \
  }

PATH_ENTRY(1)
PATH_ENTRY(2)
PATH_ENTRY(3)
