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

// One path's Whitney deposit (deposit_paths_{1,2,3}d_pyccel.py): the path, clamped to the
// grid on bounded axes, cut at every grid plane it crosses; sink(idx, axis, v) receives every
// contribution, idx[d] the edge's node index along axis d (wrapped by linear_axis).
template <int NDIM, class Sink>
__device__ __forceinline__ void whitney_path(const double *a, const double *b, double charge, const Grid &g,
                                             Sink &sink) {
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
#define PUT(i, j, k, c, v)                     \
  {                                            \
    const int64_t idx_[3] = {(i), (j), (k)};   \
    sink(idx_, (c), (v));                      \
  }
    if (NDIM == 1) {
      PUT(lo[0], 0, 0, 0, charge * (x1[0] - x0[0]));
    } else if (NDIM == 2) {
      PUT(lo[0], lo[1], 0, 0, (1.0 - f[1]) * along[0]);
      PUT(lo[0], hi[1], 0, 0, f[1] * along[0]);
      PUT(lo[0], lo[1], 0, 1, (1.0 - f[0]) * along[1]);
      PUT(hi[0], lo[1], 0, 1, f[0] * along[1]);
    } else {
      PUT(lo[0], lo[1], lo[2], 0, along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 0, 0));
      PUT(lo[0], hi[1], lo[2], 0, along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 1, 0));
      PUT(lo[0], lo[1], hi[2], 0, along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 0, 1));
      PUT(lo[0], hi[1], hi[2], 0, along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 1, 1));
      PUT(lo[0], lo[1], lo[2], 1, along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 0, 0));
      PUT(hi[0], lo[1], lo[2], 1, along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 1, 0));
      PUT(lo[0], lo[1], hi[2], 1, along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 0, 1));
      PUT(hi[0], lo[1], hi[2], 1, along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 1, 1));
      PUT(lo[0], lo[1], lo[2], 2, along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 0, 0));
      PUT(hi[0], lo[1], lo[2], 2, along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 1, 0));
      PUT(lo[0], hi[1], lo[2], 2, along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 0, 1));
      PUT(hi[0], hi[1], lo[2], 2, along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 1, 1));
    }
#undef PUT
    time = following;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) x0[d] = x1[d];
  }
}

// Whether a Whitney path's rel coordinates are usable (every one finite and below 1e300).
template <int NDIM>
__device__ __forceinline__ bool whitney_finite(const double *a, const double *b) {
  bool finite = true;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    finite = finite && a[d] > -1.0e300 && a[d] < 1.0e300 && b[d] > -1.0e300 && b[d] < 1.0e300;
  return finite;
}

template <int NDIM, int MODE>
struct GlobalSink {
  double *edges;
  unsigned long long *accum;
  double scale;
  const Grid *g;
  __device__ __forceinline__ void operator()(const int64_t *idx, int axis, double v) {
    const int64_t at = (idx[0] * g->stride[0] + (NDIM > 1 ? idx[1] * g->stride[1] : 0) +
                        (NDIM > 2 ? idx[2] * g->stride[2] : 0)) * NDIM + axis;
    deposit<MODE>(edges, accum, at, v, scale);
  }
};

template <class T, int NDIM, int MODE>
__global__ void __launch_bounds__(256) whitney_kernel(const T *__restrict__ st, int64_t ss, const T *__restrict__ en,
                                                      int64_t es, const T *__restrict__ q, int64_t n, Grid g,
                                                      double *__restrict__ edges,
                                                      unsigned long long *__restrict__ accum,
                                                      const double *__restrict__ slot) {
  GlobalSink<NDIM, MODE> sink{edges, accum, MODE == MODE_FIXED ? slot[0] : 1.0, &g};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const double charge = (double)q[p];
    if (charge == 0.0) continue;
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      a[d] = ((double)st[d * ss + p] - g.lower[d]) * g.inverse[d];
      b[d] = ((double)en[d * es + p] - g.lower[d]) * g.inverse[d];
    }
    if (!whitney_finite<NDIM>(a, b)) continue;
    whitney_path<NDIM>(a, b, charge, g, sink);
  }
}

// ---- spline edges ---------------------------------------------------------------------

__device__ __forceinline__ void gauss_rule(int count, double *point, double *weight) {
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

// The cut times of one stretch [a0, a1] (spline_edges_pyccel.py:180-230 path_cuts): the
// knot planes (shifted by `shift` for even degrees) and, on bounded axes, the walls, each
// kept when strictly inside (0, 1).  A stretch spans less than 4 cells per axis, so it
// crosses at most 4 planes per axis; every candidate has a fixed register slot (2.0 when
// absent), with the reference's arithmetic: plane = floor(low - offset) + 1 + offset,
// stepped by 1.0, t = (plane - a0) / delta.
template <int NDIM>
struct Cuts {
  static constexpr int SLOTS = NDIM * 6;
  double t[SLOTS];
};

template <int NDIM>
__device__ __forceinline__ void stretch_cuts(const double *a0, const double *a1, double shift, const Grid &g,
                                             Cuts<NDIM> &c) {
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const double delta = a1[d] - a0[d];
    const double low = delta > 0.0 ? a0[d] : a1[d];
    const double high = delta > 0.0 ? a1[d] : a0[d];
    const double offset = g.code[d] == 2 ? 0.0 : shift;
    double plane = (double)floor_int(low - offset) + 1.0 + offset;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double t = (plane - a0[d]) / delta;
      c.t[d * 6 + k] = (delta != 0.0 && plane < high && t > 0.0 && t < 1.0) ? t : 2.0;
      plane += 1.0;
    }
    double tw0 = 2.0, tw1 = 2.0;
    if (g.code[d] != 1 && delta != 0.0) {
      tw0 = (0.0 - a0[d]) / delta;
      tw1 = ((double)g.num[d] - 1.0 - a0[d]) / delta;
    }
    c.t[d * 6 + 4] = (tw0 > 0.0 && tw0 < 1.0) ? tw0 : 2.0;
    c.t[d * 6 + 5] = (tw1 > 0.0 && tw1 < 1.0) ? tw1 : 2.0;
  }
}

// The smallest cut strictly after `after` (1.0 when none): the pieces of the stretch are
// the intervals between consecutive distinct cut times, as after the reference's sort
// (equal times give the empty pieces it skips).
template <int NDIM>
__device__ __forceinline__ double next_cut(const Cuts<NDIM> &c, double after) {
  double best = 1.0;
#pragma unroll
  for (int i = 0; i < Cuts<NDIM>::SLOTS; ++i) best = (c.t[i] > after && c.t[i] < best) ? c.t[i] : best;
  return best;
}

// One path's spline-edge deposit (deposit_paths_spline_{1,2,3}d_pyccel.py; the cuts above,
// Gauss-Legendre with (p-1+(ndim-1)p)//2+1 points per piece).  sink(idx, axis, v) receives
// every contribution: idx[d] is the node (edge) index along axis d, already wrapped or
// clamped by the stencils.
template <int NDIM, int DEG, class Sink>
__device__ __forceinline__ void spline_path(const double *r0, const double *r1, double charge, const Grid &g,
                                            Sink &sink) {
  constexpr double SHIFT = DEG % 2 == 0 ? 0.5 : 0.0;
  constexpr int POINTS = (DEG - 1 + (NDIM - 1) * DEG) / 2 + 1;
  double point[5], gw[5];
  gauss_rule(POINTS, point, gw);
  double extent = 0.0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const double span = r1[d] > r0[d] ? r1[d] - r0[d] : r0[d] - r1[d];
    if (span > extent) extent = span;
  }
  const int64_t stretches = (int64_t)(extent / 4.0) + 1;
  for (int64_t st = 0; st < stretches; ++st) {
    double a0[NDIM], a1[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      a0[d] = r0[d] + (r1[d] - r0[d]) * (double)st / (double)stretches;
      a1[d] = r0[d] + (r1[d] - r0[d]) * (double)(st + 1) / (double)stretches;
    }
    Cuts<NDIM> cuts;
    stretch_cuts<NDIM>(a0, a1, SHIFT, g, cuts);
    double t0 = 0.0;
    while (t0 < 1.0) {
      const double t1 = next_cut<NDIM>(cuts, t0);
      double p0[NDIM], stp[NDIM];
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
#pragma unroll
      for (int gp = 0; gp < POINTS; ++gp) {
        double rel[NDIM];
#pragma unroll
        for (int d = 0; d < NDIM; ++d) rel[d] = p0[d] + point[gp] * stp[d];
        int64_t node[NDIM][DEG + 1];
        double wn[NDIM][DEG + 1];
        int64_t edge[NDIM][DEG > 0 ? DEG : 1];
        double we[NDIM][DEG > 0 ? DEG : 1];
        int ex[NDIM][DEG > 0 ? DEG : 1];
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          spline_axis(rel[d], g.num[d], g.code[d], DEG, node[d], wn[d]);
          along_stencil(rel[d], g.num[d], g.code[d], DEG, edge[d], we[d], ex[d]);
        }
#pragma unroll
        for (int axis = 0; axis < NDIM; ++axis) {
          const double amount = charge * gw[gp] * stp[axis];
          if (amount == 0.0) continue;
          constexpr int S = DEG + 1;
          const int s0 = axis == 0 ? DEG : S, s1 = NDIM > 1 ? (axis == 1 ? DEG : S) : 1,
                    s2 = NDIM > 2 ? (axis == 2 ? DEG : S) : 1;
#pragma unroll
          for (int a = 0; a < S; ++a) {
            if (a >= s0) continue;
            if (axis == 0 && !ex[0][a]) continue;
            const double va = amount * (axis == 0 ? we[0][a] : wn[0][a]);
            const int64_t ia = axis == 0 ? edge[0][a] : node[0][a];
#pragma unroll
            for (int b = 0; b < S; ++b) {
              if (b >= s1) continue;
              if constexpr (NDIM >= 2) {
                if (axis == 1 && !ex[1][b]) continue;
              }
              const double vb = NDIM >= 2 ? va * (axis == 1 ? we[NDIM > 1 ? 1 : 0][b] : wn[NDIM > 1 ? 1 : 0][b]) : va;
              const int64_t ib = NDIM >= 2 ? (axis == 1 ? edge[NDIM > 1 ? 1 : 0][b] : node[NDIM > 1 ? 1 : 0][b]) : 0;
#pragma unroll
              for (int e = 0; e < S; ++e) {
                if (e >= s2) continue;
                if constexpr (NDIM == 3) {
                  if (axis == 2 && !ex[2][e]) continue;
                }
                const double v = NDIM == 3 ? vb * (axis == 2 ? we[NDIM - 1][e] : wn[NDIM - 1][e]) : vb;
                const int64_t ie = NDIM == 3 ? (axis == 2 ? edge[NDIM - 1][e] : node[NDIM - 1][e]) : 0;
                const int64_t idx[3] = {ia, ib, ie};
                sink(idx, axis, v);
              }
            }
          }
        }
      }
      t0 = t1;
    }
  }
}

// The path's rel coordinates and whether they are finite.
template <class T, int NDIM>
__device__ __forceinline__ bool path_rel(const T *st, int64_t ss, const T *en, int64_t es, int64_t p, const Grid &g,
                                         double *r0, double *r1) {
  bool finite = true;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const double s = (double)st[d * ss + p], e = (double)en[d * es + p];
    finite = finite && (s - s == 0.0) && (e - e == 0.0);
    r0[d] = (s - g.lower[d]) * g.inverse[d];
    r1[d] = (e - g.lower[d]) * g.inverse[d];
  }
  return finite;
}


// One work-item per path; every contribution a global float64 atomic (MODE_ATOMIC) or an
// int64 fixed-point atomic (MODE_FIXED).
template <class T, int NDIM, int DEG, int MODE>
__global__ void __launch_bounds__(256) spline_paths_kernel(const T *__restrict__ st, int64_t ss,
                                                           const T *__restrict__ en, int64_t es,
                                                           const T *__restrict__ q, int64_t n, Grid g,
                                                           double *__restrict__ edges,
                                                           unsigned long long *__restrict__ accum,
                                                           const double *__restrict__ slot) {
  GlobalSink<NDIM, MODE> sink{edges, accum, MODE == MODE_FIXED ? slot[0] : 1.0, &g};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const double charge = (double)q[p];
    if (charge == 0.0) continue;
    double r0[NDIM], r1[NDIM];
    if (!path_rel<T, NDIM>(st, ss, en, es, p, g, r0, r1)) continue;
    spline_path<NDIM, DEG>(r0, r1, charge, g, sink);
  }
}

// ---- tiled spline-edge deposit ----------------------------------------------------------
//
// Paths binned by the cell of their start point (2D tiles of 32 x 32 cells, 3D tiles of
// 8 x 8 x 16).  One workgroup per tile accumulates the contributions of the tile's paths in
// int64 fixed point in LDS over the tile's edges plus a halo of PATH_HALO cells on every
// side (the quanta are those of the one-path-per-work-item fixed-point deposit: the same
// integers, summed exactly); a contribution outside the footprint (a long path) goes to the
// global accumulator directly.  The tile then adds its nonzero LDS words into the global
// int64 accumulator (integer atomics: exact in any order), and finalize_fixed converts.
// Each work-item walks a contiguous run of the tile's paths, so that the lanes of a wave
// work on different cells at any time (paths of one cell would hit the same LDS words).
constexpr int PATH_HALO_2D = 4, PATH_HALO_3D = 3;

template <int NDIM>
struct PathTile {
  static constexpr int TX = NDIM == 3 ? 8 : 32, TY = NDIM == 3 ? 8 : 32, TZ = NDIM == 3 ? 16 : 1;
  static constexpr int H = NDIM == 3 ? PATH_HALO_3D : PATH_HALO_2D;
  static constexpr int LX = TX + 2 * H, LY = TY + 2 * H, LZ = NDIM == 3 ? TZ + 2 * H : 1;
  static constexpr int CELLS = TX * TY * TZ;
  static constexpr int WORDS = LX * LY * LZ * NDIM;
};



// ---- one path into a tile's LDS accumulator, with int32 local indices --------------------
//
// The same arithmetic as spline_path with spline_axis / along_stencil (the rel
// wrap of a periodic axis, the clamp of a bounded one, int(r + 1/2) or int(r), the weights,
// the cut times), but node and edge indices stay unwrapped int32 numbers relative to the
// tile's footprint, folded once by +-num, and a cut time is divided out only when its
// plane or wall exists.

// The LDS accumulator as an LDS pointer (address space 3), so that its adds are LDS atomics.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

struct AxisTile {
  int num, code, base, len;  // nodes, code, first footprint node (unwrapped), footprint length
};

// The footprint slot of unwrapped node k (-1 when outside) and its wrapped or clamped node.
__device__ __forceinline__ void fold_node(int k, const AxisTile &ax, int &slot, int &node) {
  if (ax.code == 1) {
    node = k < 0 ? k + ax.num : (k >= ax.num ? k - ax.num : k);
    int l = k - ax.base;
    l = l < 0 ? l + ax.num : (l >= ax.num ? l - ax.num : l);
    slot = l < ax.len ? l : -1;
  } else {
    node = k < 0 ? 0 : (k > ax.num - 1 ? ax.num - 1 : k);
    const int l = node - ax.base;
    slot = l >= 0 && l < ax.len ? l : -1;
  }
}

__device__ __forceinline__ double wrap_rel(double r, int num) {
  return (r >= 0.0 && r < (double)num) ? r : r - (double)num * floor(r / (double)num);
}

// Node stencil of degree DEG at rel on one axis (spline_axis).
template <int DEG>
__device__ __forceinline__ void node_axis(double rel, const AxisTile &ax, int *slot, int *node, double *w) {
  double r = rel;
  if (ax.code == 1) {
    r = wrap_rel(r, ax.num);
  } else {
    if (r < 0.0) r = 0.0;
    if (r > (double)(ax.num - 1)) r = (double)ax.num - 1.0;
  }
  const int middle = (DEG % 2 == 0) ? (int)(r + 0.5) : (int)r;
  const double o = r - (double)middle;
  if (DEG == 1) {
    w[0] = 1.0 - o;
    w[1] = o;
  } else if (DEG == 2) {
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
  for (int a = 0; a <= DEG; ++a) fold_node(middle - DEG / 2 + a, ax, slot[a], node[a]);
}

// Edge stencil of degree DEG at rel on one axis (along_stencil): DEG edges, weights
// B_{DEG-1}(rel - k - 1/2); an edge past a bounded wall does not exist (slot -2).
template <int DEG>
__device__ __forceinline__ void edge_axis(double rel, const AxisTile &ax, int *slot, int *edge, double *w) {
  constexpr int LOWER = DEG - 1;
  double r = rel;
  if (ax.code == 1) r = wrap_rel(r, ax.num);
  const double shifted = r - 0.5;
  int middle;
  if (LOWER % 2 == 0) {
    middle = (int)floor_int(shifted + 0.5);
    if (ax.code == 0 && middle > ax.num - 2) middle = ax.num - 2;
  } else {
    middle = (int)floor_int(shifted);
  }
  const double o = shifted - (double)middle;
  if (LOWER == 0) {
    w[0] = 1.0;
  } else if (LOWER == 1) {
    w[0] = 1.0 - o;
    w[1] = o;
  } else {
    w[0] = 0.5 * (0.5 - o) * (0.5 - o);
    w[1] = 0.75 - o * o;
    w[2] = 0.5 * (0.5 + o) * (0.5 + o);
  }
#pragma unroll
  for (int a = 0; a < DEG; ++a) {
    const int k = middle - LOWER / 2 + a;
    if (ax.code == 1) {
      fold_node(k, ax, slot[a], edge[a]);
    } else if (k < 0 || k > ax.num - 2) {
      slot[a] = -2;
      edge[a] = 0;
    } else {
      edge[a] = k;
      const int l = k - ax.base;
      slot[a] = l >= 0 && l < ax.len ? l : -1;
    }
  }
}

// ---- spline-edge paths into a tile, piece by piece ---------------------------------------
//
// The walk of spline_path (stretches of at most four cells, cut at the knot planes and the
// walls), with the work of one piece at a time: between two consecutive cuts no knot plane
// is crossed, so every Gauss point of a piece has the same node and edge stencil.  A piece
// takes its stencil once (the nodes and edges around its midpoint, unwrapped, folded into
// the footprint once), sums the Gauss points' contributions per edge in float64 registers
// and adds each sum into the tile as one fixed-point word (to_fixed_small: the automatic
// scale keeps a path's whole charge times its length below 2^49 quanta).  The offsets of a
// point from the piece's middle node and edge are r - middle in unwrapped coordinates,
// equal to the reference's wrapped r - middle.

// The state of one work-item's current path: its stretch, the stretch's cut times and the
// start of the next piece.
template <int NDIM>
struct PathWalk {
  double r0[NDIM], r1[NDIM], a0[NDIM], a1[NDIM], cut[NDIM * 6];
  double charge, t0;
  int st, stretches;
};

// The cut times of stretch w.st (a slot per candidate plane or wall, divided out only when
// it exists; 2.0 when absent) and its first piece.
template <int NDIM, int DEG>
__device__ __forceinline__ void walk_stretch(PathWalk<NDIM> &w, const Grid &g) {
  constexpr double SHIFT = DEG % 2 == 0 ? 0.5 : 0.0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    double f0 = (w.r1[d] - w.r0[d]) * (double)w.st, f1 = (w.r1[d] - w.r0[d]) * (double)(w.st + 1);
    if (w.stretches != 1) {  // a division by 1.0 is exact
      f0 = f0 / (double)w.stretches;
      f1 = f1 / (double)w.stretches;
    }
    w.a0[d] = w.r0[d] + f0;
    w.a1[d] = w.r0[d] + f1;
    const double delta = w.a1[d] - w.a0[d];
    const double low = delta > 0.0 ? w.a0[d] : w.a1[d];
    const double high = delta > 0.0 ? w.a1[d] : w.a0[d];
    const double offset = g.code[d] == 2 ? 0.0 : SHIFT;
    double plane = (double)floor_int(low - offset) + 1.0 + offset;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double t = 2.0;
      if (delta != 0.0 && plane < high) {
        t = (plane - w.a0[d]) / delta;
        t = t > 0.0 && t < 1.0 ? t : 2.0;
      }
      w.cut[d * 6 + k] = t;
      plane += 1.0;
    }
    double tw0 = 2.0, tw1 = 2.0;
    if (g.code[d] != 1 && delta != 0.0 && (low < 0.0 || high > (double)g.num[d] - 1.0)) {
      tw0 = (0.0 - w.a0[d]) / delta;
      tw1 = ((double)g.num[d] - 1.0 - w.a0[d]) / delta;
      tw0 = tw0 > 0.0 && tw0 < 1.0 ? tw0 : 2.0;
      tw1 = tw1 > 0.0 && tw1 < 1.0 ? tw1 : 2.0;
    }
    w.cut[d * 6 + 4] = tw0;
    w.cut[d * 6 + 5] = tw1;
  }
  w.t0 = 0.0;
}

// The node and edge stencil of a piece on one axis from its midpoint rm (clamped to the grid
// on a bounded axis; on a periodic axis taken modulo num, so that the middle node is in
// [0, num] and one fold reaches every index): the offsets of the midpoint from the middle
// node and the middle edge (a Gauss point's are these plus its distance from the midpoint),
// and the footprint slots and indices of the stencil (node_axis, edge_axis).
template <int DEG>
struct PieceAxis {
  double dn, de;
  int ns[DEG + 1], nn[DEG + 1], es[DEG], en[DEG];
};

template <int DEG>
__device__ __forceinline__ void piece_axis(double rm, const AxisTile &ax, PieceAxis<DEG> &pa) {
  constexpr int LOWER = DEG - 1;
  const double r = ax.code == 1 ? wrap_rel(rm, ax.num) : rm;
  const int nmid = (int)((DEG % 2 == 0) ? floor_int(r + 0.5) : floor_int(r));
  pa.dn = r - (double)nmid;
#pragma unroll
  for (int a = 0; a <= DEG; ++a) fold_node(nmid - DEG / 2 + a, ax, pa.ns[a], pa.nn[a]);
  const double shifted = r - 0.5;
  int emid;
  if (LOWER % 2 == 0) {
    emid = (int)floor_int(shifted + 0.5);
    if (ax.code == 0 && emid > ax.num - 2) emid = ax.num - 2;
  } else {
    emid = (int)floor_int(shifted);
  }
  pa.de = shifted - (double)emid;
#pragma unroll
  for (int a = 0; a < DEG; ++a) {
    const int k = emid - LOWER / 2 + a;
    if (ax.code == 1) {
      fold_node(k, ax, pa.es[a], pa.en[a]);
    } else if (k < 0 || k > ax.num - 2) {
      pa.es[a] = -2;
      pa.en[a] = 0;
    } else {
      pa.en[a] = k;
      const int l = k - ax.base;
      pa.es[a] = l >= 0 && l < ax.len ? l : -1;
    }
  }
}

template <int P>
__device__ __forceinline__ void bspline_weights(double o, double *w) {
  if (P == 0) {
    w[0] = 1.0;
  } else if (P == 1) {
    w[0] = 1.0 - o;
    w[1] = o;
  } else if (P == 2) {
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
}

// One contribution of a piece into the tile (or the global accumulator outside it).
template <int NDIM>
__device__ __forceinline__ void piece_add(double v, int axis, const int (&slot)[3], const int (&idx)[3],
                                          const Grid &g, lds_u64 *lds, unsigned long long *__restrict__ accum,
                                          double scale) {
  using PT = PathTile<NDIM>;
  if (slot[0] == -2 || slot[1] == -2 || slot[2] == -2) return;  // an edge past a bounded wall
  const unsigned long long f = to_fixed_small(v * scale);
  if (f == 0ull) return;
  if (slot[0] >= 0 && slot[1] >= 0 && slot[2] >= 0) {
    __hip_atomic_fetch_add(lds + ((slot[0] * PT::LY + slot[1]) * PT::LZ + slot[2]) * NDIM + axis, f,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    const int64_t at = ((int64_t)idx[0] * g.stride[0] + (NDIM > 1 ? (int64_t)idx[1] * g.stride[1] : 0) +
                        (NDIM > 2 ? (int64_t)idx[2] * g.stride[2] : 0)) * NDIM + axis;
    atomicAdd(accum + at, f);
  }
}

// The piece [w.t0, t1) of the current stretch into the tile.
template <int NDIM, int DEG>
__device__ __forceinline__ void piece_into_tile(const PathWalk<NDIM> &w, double t1, const Grid &g,
                                                const AxisTile *ax, lds_u64 *lds,
                                                unsigned long long *__restrict__ accum, double scale) {
  constexpr int POINTS = (DEG - 1 + (NDIM - 1) * DEG) / 2 + 1;
  constexpr int S = DEG + 1;
  double point[5], gw[5];
  gauss_rule(POINTS, point, gw);
  double p0[NDIM], stp[NDIM];
  PieceAxis<DEG> pa[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    double begin = w.a0[d] + w.t0 * (w.a1[d] - w.a0[d]);
    double finish = w.a0[d] + t1 * (w.a1[d] - w.a0[d]);
    if (g.code[d] != 1) {
      const double top = (double)(g.num[d] - 1);
      if (begin < 0.0) begin = 0.0;
      if (begin > top) begin = (double)g.num[d] - 1.0;
      if (finish < 0.0) finish = 0.0;
      if (finish > top) finish = (double)g.num[d] - 1.0;
    }
    p0[d] = begin;
    stp[d] = finish - begin;
    piece_axis<DEG>(begin + 0.5 * stp[d], ax[d], pa[d]);
  }
  // per axis the sums over the Gauss points of charge * gw * step * (edge weight along the
  // axis) * (node weights across), indexed [a][b][e] as the stencils
  constexpr int NB = NDIM > 1 ? S : 1, NE = NDIM > 2 ? S : 1;
  double sum[NDIM][S][NB][NE];
#pragma unroll
  for (int axis = 0; axis < NDIM; ++axis)
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int e = 0; e < NE; ++e) sum[axis][a][b][e] = 0.0;
#pragma unroll
  for (int gp = 0; gp < POINTS; ++gp) {
    double wn[NDIM][S], we[NDIM][DEG];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const double off = (point[gp] - 0.5) * stp[d];  // from the midpoint
      bspline_weights<DEG>(pa[d].dn + off, wn[d]);
      bspline_weights<DEG - 1>(pa[d].de + off, we[d]);
    }
#pragma unroll
    for (int axis = 0; axis < NDIM; ++axis) {
      const double amount = w.charge * gw[gp] * stp[axis];
#pragma unroll
      for (int a = 0; a < (axis == 0 ? DEG : S); ++a) {
        const double va = amount * (axis == 0 ? we[0][a] : wn[0][a]);
#pragma unroll
        for (int b = 0; b < (NDIM < 2 ? 1 : (axis == 1 ? DEG : S)); ++b) {
          const double vb = NDIM < 2 ? va : va * (axis == 1 ? we[NDIM > 1 ? 1 : 0][b] : wn[NDIM > 1 ? 1 : 0][b]);
#pragma unroll
          for (int e = 0; e < (NDIM < 3 ? 1 : (axis == 2 ? DEG : S)); ++e) {
            const double v = NDIM < 3 ? vb : vb * (axis == 2 ? we[NDIM - 1][e] : wn[NDIM - 1][e]);
            sum[axis][a][b][e] += v;
          }
        }
      }
    }
  }
#pragma unroll
  for (int axis = 0; axis < NDIM; ++axis) {
    if (stp[axis] == 0.0) continue;  // no motion along this axis in this piece
#pragma unroll
    for (int a = 0; a < (axis == 0 ? DEG : S); ++a)
#pragma unroll
      for (int b = 0; b < (NDIM < 2 ? 1 : (axis == 1 ? DEG : S)); ++b)
#pragma unroll
        for (int e = 0; e < (NDIM < 3 ? 1 : (axis == 2 ? DEG : S)); ++e) {
          const int slot[3] = {axis == 0 ? pa[0].es[a] : pa[0].ns[a],
                               NDIM < 2 ? 0 : (axis == 1 ? pa[NDIM > 1 ? 1 : 0].es[b] : pa[NDIM > 1 ? 1 : 0].ns[b]),
                               NDIM < 3 ? 0 : (axis == 2 ? pa[NDIM - 1].es[e] : pa[NDIM - 1].ns[e])};
          const int idx[3] = {axis == 0 ? pa[0].en[a] : pa[0].nn[a],
                              NDIM < 2 ? 0 : (axis == 1 ? pa[NDIM > 1 ? 1 : 0].en[b] : pa[NDIM > 1 ? 1 : 0].nn[b]),
                              NDIM < 3 ? 0 : (axis == 2 ? pa[NDIM - 1].en[e] : pa[NDIM - 1].nn[e])};
          piece_add<NDIM>(sum[axis][a][b][e], axis, slot, idx, g, lds, accum, scale);
        }
  }
}

// The next piece's end: the smallest cut strictly after t0, 1.0 when none.
template <int NDIM>
__device__ __forceinline__ double piece_end(const PathWalk<NDIM> &w) {
  double t1 = 1.0;
#pragma unroll
  for (int i = 0; i < NDIM * 6; ++i) t1 = (w.cut[i] > w.t0 && w.cut[i] < t1) ? w.cut[i] : t1;
  return t1;
}

// Whitney contributions into a tile's LDS accumulator: the wrapped node index of each axis
// folded into the footprint (one +-num step), a contribution outside it to the global
// accumulator.  The tiles run under the automatic scale only, which keeps every
// contribution below 2^49 quanta (|q| times the path length bounds it), so the rounding is
// to_fixed_small's.
template <int NDIM>
struct WhitneyTileSink {
  lds_u64 *lds;
  unsigned long long *accum;
  double scale;
  const Grid *g;
  const AxisTile *ax;
  __device__ __forceinline__ void operator()(const int64_t *idx, int axis, double v) {
    using PT = PathTile<NDIM>;
    int l[3] = {0, 0, 0};
    bool in = true;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      int k = (int)idx[d] - ax[d].base;
      if (ax[d].code == 1) k = k < 0 ? k + ax[d].num : (k >= ax[d].num ? k - ax[d].num : k);
      in = in && k >= 0 && k < ax[d].len;
      l[d] = k;
    }
    const unsigned long long f = to_fixed_small(v * scale);
    if (in) {
      __hip_atomic_fetch_add(lds + ((l[0] * PT::LY + l[1]) * PT::LZ + l[2]) * NDIM + axis, f, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      const int64_t at = (idx[0] * g->stride[0] + (NDIM > 1 ? idx[1] * g->stride[1] : 0) +
                          (NDIM > 2 ? idx[2] * g->stride[2] : 0)) * NDIM + axis;
      atomicAdd(accum + at, f);
    }
  }
};

// DEG 1..3: spline edges (piece by piece, one piece per trip); DEG 0: Whitney.  2D tiles:
// four workgroups of 256 per CU, at most 128 VGPRs (four waves per SIMD: the work-items walk
// long chains of dependent float64 arithmetic, and more waves hide more of it).  A 3D
// tile's LDS (103 KB) leaves room for one workgroup per CU: 1024 work-items for the
// Whitney walk (2^23 paths on 128^3: 2.7 -> 1.9 ms against 256), 512 for the spline edges,
// whose per-piece sums hold up to 3 x 18 float64 registers.
template <class T, int NDIM, int DEG, int BLOCK>
__global__ void __launch_bounds__(BLOCK, BLOCK == 256 ? 4 : 1) spline_paths_tiled(const T *__restrict__ st,
                                                            const T *__restrict__ q, const T *__restrict__ en,
                                                            int64_t es, int64_t n, const uint32_t *__restrict__ start, Grid g,
                                                            int nty, int ntz, unsigned long long *__restrict__ accum,
                                                            const double *__restrict__ slot) {
  using PT = PathTile<NDIM>;
  __shared__ unsigned long long lds[PT::WORDS];
  const int tile = blockIdx.x, t = threadIdx.x;
  const int tz = NDIM == 3 ? tile % ntz : 0, ty = NDIM == 3 ? (tile / ntz) % nty : tile % nty;
  const int tx = NDIM == 3 ? tile / (ntz * nty) : tile / nty;
  const int64_t base[3] = {(int64_t)tx * PT::TX - PT::H, (int64_t)ty * PT::TY - PT::H, (int64_t)tz * PT::TZ - PT::H};
  const double scale = slot[0];
  AxisTile ax[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    ax[d] = AxisTile{(int)g.num[d], g.code[d], (int)base[d], d == 0 ? PT::LX : (d == 1 ? PT::LY : PT::LZ)};
  for (int i = t; i < PT::WORDS; i += BLOCK) lds[i] = 0ull;
  const uint32_t first = start[(int64_t)tile * PT::CELLS], last = start[(int64_t)(tile + 1) * PT::CELLS];
  __syncthreads();
  // 2D tiles hold tens of thousands of paths: lanes take runs of 8 consecutive paths, run t
  // of each pass of BLOCK runs, so a wavefront's row loads stay coalesced (a lane's share as
  // one run made every load touch 64 lines: Whitney 5.5 -> 3.4 ms for 2^25 paths on 1024^2).
  // 3D tiles hold a few thousand for 1024 lanes: an even split, one run per lane.
  const uint32_t count = last - first;
  const uint32_t run = NDIM == 2 ? 8u : (count > 0 ? (count + BLOCK - 1) / BLOCK : 1u);
  if constexpr (DEG == 0) {
    WhitneyTileSink<NDIM> wsink{(lds_u64 *)lds, accum, scale, &g, ax};
    for (uint32_t lo = first + t * run; lo < last; lo += BLOCK * run)
      for (uint32_t p = lo; p < (lo + run < last ? lo + run : last); ++p) {
        const double charge = (double)q[p];
        if (charge == 0.0) continue;
        double r0[NDIM], r1[NDIM];
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          r0[d] = ((double)st[d * n + p] - g.lower[d]) * g.inverse[d];
          r1[d] = ((double)en[d * es + p] - g.lower[d]) * g.inverse[d];
        }
        if (!whitney_finite<NDIM>(r0, r1)) continue;
        whitney_path<NDIM>(r0, r1, charge, g, wsink);
      }
  } else {
    // One piece per trip: a work-item whose path is done takes its next live path (and its
    // first stretch) in the same trip, so the lanes of a wavefront stay busy while their
    // paths have different numbers of pieces.  The k-th path of work-item t is path
    // first + (k / run) * BLOCK * run + t * run + k % run.
    uint32_t k = 0;
    PathWalk<NDIM> w;
    bool live = false;
    for (;;) {
      while (!live) {
        const uint32_t p = first + (k / run) * (BLOCK * run) + t * run + k % run;
        if (p >= last) break;  // no paths left (later ones lie further on)
        ++k;
        const double charge = (double)q[p];
        bool finite = true;
        double extent = 0.0;
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          const double s0 = (double)st[d * n + p], e0 = (double)en[d * es + p];
          finite = finite && (s0 - s0 == 0.0) && (e0 - e0 == 0.0);
          w.r0[d] = (s0 - g.lower[d]) * g.inverse[d];
          w.r1[d] = (e0 - g.lower[d]) * g.inverse[d];
          const double span = w.r1[d] > w.r0[d] ? w.r1[d] - w.r0[d] : w.r0[d] - w.r1[d];
          extent = span > extent ? span : extent;
        }
        if (charge == 0.0 || !finite) continue;
        w.charge = charge;
        w.stretches = (int)(extent / 4.0) + 1;
        w.st = 0;
        walk_stretch<NDIM, DEG>(w, g);
        live = true;
      }
      if (!live) break;
      const double t1 = piece_end<NDIM>(w);
      piece_into_tile<NDIM, DEG>(w, t1, g, ax, (lds_u64 *)lds, accum, scale);
      w.t0 = t1;
      if (t1 >= 1.0) {
        if (++w.st < w.stretches) walk_stretch<NDIM, DEG>(w, g);
        else live = false;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < PT::WORDS; i += BLOCK) {
    const unsigned long long v = lds[i];
    if (v == 0ull) continue;
    const int axis = i % NDIM, node = i / NDIM;
    const int l[3] = {NDIM == 3 ? node / (PT::LY * PT::LZ) : node / PT::LY,
                      NDIM == 3 ? (node / PT::LZ) % PT::LY : node % PT::LY, NDIM == 3 ? node % PT::LZ : 0};
    int64_t at = 0;
    bool inside = true;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      int64_t k = base[d] + l[d];
      if (g.code[d] == 1 && (k < 0 || k >= g.num[d])) k = ((k % g.num[d]) + g.num[d]) % g.num[d];
      inside = inside && k >= 0 && k < g.extent[d];
      at += k * g.stride[d];
    }
    if (inside) atomicAdd(accum + at * NDIM + axis, v);
  }
}

// The fixed-point bound of a path deposit: every edge value is at most
// sum_paths |q| * (length in cells + 1).
template <class T>
__global__ void path_bound_kernel(const T *__restrict__ st, int64_t ss, const T *__restrict__ en, int64_t es,
                                  const T *__restrict__ q, int64_t n, Grid g, unsigned long long *__restrict__ out) {
  double m = 0.0;
  // two paths of the grid-stride walk per trip, their loads issued together
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += 2 * stride) {
    const int64_t p2 = p + stride < n ? p + stride : p;
    double len = 0.0, len2 = 0.0;
    for (int d = 0; d < g.ndim; ++d) {
      len += fabs(((double)en[d * es + p] - (double)st[d * ss + p]) * g.inverse[d]);
      len2 += fabs(((double)en[d * es + p2] - (double)st[d * ss + p2]) * g.inverse[d]);
    }
    const double b = fabs((double)q[p]) * (len + 2.0), b2 = fabs((double)q[p2]) * (len2 + 2.0);
    if (b > m && b < 1.0e300) m = b;
    if (b2 > m && b2 < 1.0e300) m = b2;
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

template <class T, int NDIM, int DEG>
static void launch_spline_paths(int mode, const void *st, int64_t ss, const void *en, int64_t es, const void *q,
                                int64_t n, const Grid &g, double *edges, unsigned long long *accum,
                                const double *slot, hipStream_t stream) {
  const int threads = 256, blocks = blocks_for(n, threads, 1 << 16);
  if (mode == MODE_ATOMIC)
    hipLaunchKernelGGL((spline_paths_kernel<T, NDIM, DEG, MODE_ATOMIC>), dim3(blocks), dim3(threads), 0, stream,
                       (const T *)st, ss, (const T *)en, es, (const T *)q, n, g, edges, accum, slot);
  else
    hipLaunchKernelGGL((spline_paths_kernel<T, NDIM, DEG, MODE_FIXED>), dim3(blocks), dim3(threads), 0, stream,
                       (const T *)st, ss, (const T *)en, es, (const T *)q, n, g, edges, accum, slot);
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
  } else if (degree == 1) {
    launch_spline_paths<T, NDIM, 1>(mode, st, ss, en, es, q, n, g, edges, accum, slot, stream);
  } else if (degree == 2) {
    launch_spline_paths<T, NDIM, 2>(mode, st, ss, en, es, q, n, g, edges, accum, slot, stream);
  } else {
    launch_spline_paths<T, NDIM, 3>(mode, st, ss, en, es, q, n, g, edges, accum, slot, stream);
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

template <class T, int NDIM>
static void launch_paths_tiled(int degree, const void *spos, const void *q, const void *en, int64_t es, int64_t n,
                               const uint32_t *bins,
                               const Grid &g, unsigned long long *accum, const double *slot, hipStream_t stream) {
  using PT = PathTile<NDIM>;
  const int ntx = (int)((g.num[0] + PT::TX - 1) / PT::TX), nty = (int)((g.num[1] + PT::TY - 1) / PT::TY);
  const int ntz = NDIM == 3 ? (int)((g.num[2] + PT::TZ - 1) / PT::TZ) : 1;
  const dim3 grid(ntx * nty * ntz);
  ProfileScope prof("spline_paths_tiled", stream);
  // 3D: 1024 work-items for the Whitney walk (125 VGPRs), 512 for the spline edges, whose
  // per-piece sums need more than the 128 VGPRs of 1024 (2^23 paths on 128^3, degree 2:
  // 512 work-items without spills 2.8 ms, 1024 with 120 VGPRs spilled 4.7 ms)
#define TILED(D)                                                                                                   \
  hipLaunchKernelGGL((spline_paths_tiled<T, NDIM, D, NDIM == 3 ? (D == 0 ? 1024 : 512) : 256>), grid,             \
                     dim3(NDIM == 3 ? (D == 0 ? 1024 : 512) : 256), 0, stream,                                     \
                     (const T *)spos, (const T *)q, (const T *)en, es, n, bins, g, nty, ntz, accum, slot)
  if (degree == 0) TILED(0);  // Whitney
  else if (degree == 1) TILED(1);
  else if (degree == 2) TILED(2);
  else TILED(3);
#undef TILED
}

//: paths at least this many go through the tiled deposit (spline edges, fixed point, 2D/3D)
constexpr int64_t PATHS_TILED_MIN = 1 << 16;

// Bin the paths by start cell for the tiled deposit: positions = start, values = the
// charge and the end point (so that all of a path moves with it); returns the bins' sorted
// start points, sorted rows (charge, end) and offsets in ctx->path_sorted.
static int bin_paths(int kind, int degree, int ndim, const void *start, int64_t ss, const void *end, int64_t es,
                     const void *charges, int dtype, int64_t n, const double *lower, const double *cell,
                     const int64_t *num, const int64_t *code, mpf_context *ctx, hipStream_t stream, void **spos,
                     void **srows, uint32_t **bins) {
  const size_t elem = dtype == F32 ? 4 : 8, nn = (size_t)n;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  int rc = ensure(ctx->path_in, up(nn * elem * (ndim + 1)), false, stream);
  if (rc) return rc;
  char *in = (char *)ctx->path_in.ptr;
  MPF_CHECK(hipMemcpyAsync(in, charges, nn * elem, hipMemcpyDeviceToDevice, stream));
  MPF_CHECK(hipMemcpy2DAsync(in + nn * elem, nn * elem, end, (size_t)es * elem, nn * elem, (size_t)ndim,
                             hipMemcpyDeviceToDevice, stream));
  int64_t cells = 0;
  rc = mpf_bin_sizes(ndim, num, &cells);
  if (rc) return rc;
  const size_t b_spos = up(nn * elem * ndim), b_rows = up(nn * elem * (ndim + 1)), b_order = up(nn * 8),
               b_bins = up((size_t)(cells + 2) * 4);
  rc = ensure(ctx->path_sorted, b_spos + b_rows + b_order + b_bins + 256, false, stream);
  if (rc) return rc;
  char *out = (char *)ctx->path_sorted.ptr;
  *spos = out;
  *srows = out + b_spos;
  int64_t *order = (int64_t *)(out + b_spos + b_rows);
  *bins = (uint32_t *)(out + b_spos + b_rows + b_order);
  double *scalars = (double *)(out + b_spos + b_rows + b_order + b_bins);
  return mpf_bin_markers(kind, degree, ndim, start, ss, in, (int64_t)nn, dtype, n, ndim + 1, lower, cell, num, code,
                         *spos, *srows, order, *bins, scalars, MPF_SCALE_AUTO, ctx, (mpf_stream)stream);
}

// ---- edge gathers (the transpose of the path deposits) ----------------------------------
//
// One marker per work-item.  DEG 0: WHITNEY (gather_edges_{1,2,3}d_pyccel.py): in LINEAR's
// cell of the marker the edges along axis a weighted by the linear weights across, the
// products and sums in the reference's order, over the cell size along a.  DEG 1..3: the
// spline edges (gather_edges_spline_{1,2,3}d_pyccel.py): the position clamped to the grid
// on a non-periodic axis; along axis a the along_stencil edges that exist times the
// spline_axis nodes across.
template <class T, int NDIM, int DEG>
__global__ void __launch_bounds__(256) gather_edges_kernel(const T *__restrict__ pos, int64_t ps, int64_t n, Grid g,
                                                           const double *__restrict__ edges, T *__restrict__ res,
                                                           int64_t rs) {
  const int64_t s0 = NDIM == 3 ? g.extent[1] * g.extent[2] : (NDIM == 2 ? g.extent[1] : 1);
  const int64_t s1 = NDIM == 3 ? g.extent[2] : 1;
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    if constexpr (DEG == 0) {
      int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      double f[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        int64_t nd[2];
        double wd[2];
        linear_axis(((double)pos[d * ps + m] - g.lower[d]) * g.inverse[d], g.num[d], g.code[d], nd, wd);
        lo[d] = nd[0];
        hi[d] = nd[1];
        f[d] = wd[1];
      }
      if constexpr (NDIM == 1) {
        res[m] = (T)(edges[lo[0]] * g.inverse[0]);
      } else if constexpr (NDIM == 2) {
#define E2(i, j, a) edges[((i) * s0 + (j)) * 2 + (a)]
        res[m] = (T)(((1.0 - f[1]) * E2(lo[0], lo[1], 0) + f[1] * E2(lo[0], hi[1], 0)) * g.inverse[0]);
        res[rs + m] = (T)(((1.0 - f[0]) * E2(lo[0], lo[1], 1) + f[0] * E2(hi[0], lo[1], 1)) * g.inverse[1]);
#undef E2
      } else {
#define E3(i, j, k, a) edges[((i) * s0 + (j) * s1 + (k)) * 3 + (a)]
        const double wx0 = 1.0 - f[0], wx1 = f[0], wy0 = 1.0 - f[1], wy1 = f[1], wz0 = 1.0 - f[2], wz1 = f[2];
        res[m] = (T)((wy0 * wz0 * E3(lo[0], lo[1], lo[2], 0) + wy1 * wz0 * E3(lo[0], hi[1], lo[2], 0) +
                      wy0 * wz1 * E3(lo[0], lo[1], hi[2], 0) + wy1 * wz1 * E3(lo[0], hi[1], hi[2], 0)) * g.inverse[0]);
        res[rs + m] = (T)((wx0 * wz0 * E3(lo[0], lo[1], lo[2], 1) + wx1 * wz0 * E3(hi[0], lo[1], lo[2], 1) +
                           wx0 * wz1 * E3(lo[0], lo[1], hi[2], 1) + wx1 * wz1 * E3(hi[0], lo[1], hi[2], 1)) *
                          g.inverse[1]);
        res[2 * rs + m] = (T)((wx0 * wy0 * E3(lo[0], lo[1], lo[2], 2) + wx1 * wy0 * E3(hi[0], lo[1], lo[2], 2) +
                               wx0 * wy1 * E3(lo[0], hi[1], lo[2], 2) + wx1 * wy1 * E3(hi[0], hi[1], lo[2], 2)) *
                              g.inverse[2]);
#undef E3
      }
    } else {
      constexpr int S = DEG + 1;
      double rel[NDIM];
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        double r = ((double)pos[d * ps + m] - g.lower[d]) * g.inverse[d];
        if (g.code[d] != 1) {
          if (r < 0.0) r = 0.0;
          if (r > (double)(g.num[d] - 1)) r = (double)g.num[d] - 1.0;
        }
        rel[d] = r;
      }
      int64_t node[NDIM][S], edge[NDIM][DEG];
      double wn[NDIM][S], we[NDIM][DEG];
      int ex[NDIM][DEG];
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        spline_axis(rel[d], g.num[d], g.code[d], DEG, node[d], wn[d]);
        along_stencil(rel[d], g.num[d], g.code[d], DEG, edge[d], we[d], ex[d]);
      }
#pragma unroll
      for (int axis = 0; axis < NDIM; ++axis) {
        double total = 0.0;
#pragma unroll
        for (int a = 0; a < (axis == 0 ? DEG : S); ++a) {
          if (axis == 0 && !ex[0][a]) continue;
          const double va = axis == 0 ? we[0][a] : wn[0][a];
          const int64_t ia = axis == 0 ? edge[0][a] : node[0][a];
#pragma unroll
          for (int b = 0; b < (NDIM < 2 ? 1 : (axis == 1 ? DEG : S)); ++b) {
            if (NDIM >= 2 && axis == 1 && !ex[NDIM > 1 ? 1 : 0][b]) continue;
            const double vb = NDIM < 2 ? va : va * (axis == 1 ? we[NDIM > 1 ? 1 : 0][b] : wn[NDIM > 1 ? 1 : 0][b]);
            const int64_t ib = NDIM < 2 ? 0 : (axis == 1 ? edge[NDIM > 1 ? 1 : 0][b] : node[NDIM > 1 ? 1 : 0][b]);
#pragma unroll
            for (int e = 0; e < (NDIM < 3 ? 1 : (axis == 2 ? DEG : S)); ++e) {
              if (NDIM == 3 && axis == 2 && !ex[NDIM - 1][e]) continue;
              const double v = NDIM < 3 ? vb : vb * (axis == 2 ? we[NDIM - 1][e] : wn[NDIM - 1][e]);
              const int64_t ie = NDIM < 3 ? 0 : (axis == 2 ? edge[NDIM - 1][e] : node[NDIM - 1][e]);
              total += v * edges[(ia * s0 + ib * s1 + ie) * NDIM + axis];
            }
          }
        }
        res[axis * rs + m] = (T)(total * g.inverse[axis]);
      }
    }
  }
}

template <class T, int NDIM>
static void launch_gather_edges(int degree, const void *pos, int64_t ps, int64_t n, const Grid &g,
                                const double *edges, void *res, int64_t rs, hipStream_t stream) {
  const int threads = 256, blocks = blocks_for(n, threads, 1 << 16);
#define GE(D)                                                                                                    \
  hipLaunchKernelGGL((gather_edges_kernel<T, NDIM, D>), dim3(blocks), dim3(threads), 0, stream, (const T *)pos, ps, \
                     n, g, edges, (T *)res, rs)
  if (degree == 0) GE(0);
  else if (degree == 1) GE(1);
  else if (degree == 2) GE(2);
  else GE(3);
#undef GE
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
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  const int64_t count = g.extent[0] * (ndim > 1 ? g.extent[1] : 1) * (ndim > 2 ? g.extent[2] : 1) * ndim;
  unsigned long long *accum = nullptr;
  const double *slot = nullptr;
  // large spline-edge deposits in fixed point: binned by start cell, tiled (the same sums)
  // (with the automatic scale, under which every contribution stays below 2^49 quanta and
  // the tiles round it with to_fixed_small)
  bool tiled = mode == MODE_FIXED && (ndim == 2 || ndim == 3) && n >= PATHS_TILED_MIN && scale_exp == MPF_SCALE_AUTO;
  for (int d = 0; d < ndim; ++d) tiled = tiled && code[d] != 2;
  void *spos = nullptr, *srows = nullptr;
  uint32_t *bins = nullptr;
  if (tiled) {
    // the binning's own scale goes through ctx->scalars first; the path scale follows
    // Whitney paths by the linear cell of their start, spline-edge paths by the spline cell
    rc = bin_paths(kind == 0 ? LINEAR : SPLINE, deg, ndim, start, start_stride, end, end_stride, charges, dtype, n,
                   lower, cell_size, num_nodes, code, ctx, stream, &spos, &srows, &bins);
    if (rc) return rc;
  }
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
  if (tiled) {
    const int tdeg = kind == 0 ? 0 : deg;  // degree 0 selects the Whitney tile kernel
    // the sorted rows: the charge, then the end point
    const size_t elem = dtype == F32 ? 4 : 8;
    const void *sq = srows, *sen = (const char *)srows + (size_t)n * elem;
    if (ndim == 2 && dtype == F32) launch_paths_tiled<float, 2>(tdeg, spos, sq, sen, n, n, bins, g, accum, slot, stream);
    if (ndim == 2 && dtype == F64) launch_paths_tiled<double, 2>(tdeg, spos, sq, sen, n, n, bins, g, accum, slot, stream);
    if (ndim == 3 && dtype == F32) launch_paths_tiled<float, 3>(tdeg, spos, sq, sen, n, n, bins, g, accum, slot, stream);
    if (ndim == 3 && dtype == F64) launch_paths_tiled<double, 3>(tdeg, spos, sq, sen, n, n, bins, g, accum, slot, stream);
  } else if (dtype == F32) {
    dispatch_paths<float>(kind, ndim, mode, start, start_stride, end, end_stride, charges, n, g, deg, edges, accum,
                          slot, stream);
  } else {
    dispatch_paths<double>(kind, ndim, mode, start, start_stride, end, end_stride, charges, n, g, deg, edges, accum,
                           slot, stream);
  }
  MPF_LAUNCH_CHECK();
  if (mode == MODE_FIXED) return finalize_fixed(ctx, edges, count, stream);
  return MPF_OK;
}

// Paths kept binned by the caller (sorted_start in the bins' order, with end and charges in
// the same order): the tiled fixed-point deposit without the sort.  Whitney paths read bins
// of the linear cell, spline-edge paths bins of the spline nodes of their degree.
extern "C" int mpf_deposit_paths_binned(int kind, int degree, int ndim, const void *sorted_start, const void *end,
                                        int64_t end_stride, const void *charges, int dtype, int64_t n,
                                        const uint32_t *start, double *edges, const double *lower,
                                        const double *cell_size, const int64_t *num_nodes, const int64_t *code,
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
  if (ndim != 2 && ndim != 3) {
    set_error("binned path deposits are 2D or 3D, got %dD", ndim);
    return MPF_ERR_UNSUPPORTED;
  }
  if (n < 0 || (n > 0 && end_stride < n)) {
    set_error("need n >= 0 and end_stride >= n");
    return MPF_ERR_ARGUMENT;
  }
  for (int d = 0; d < ndim; ++d)
    if (code[d] != 0 && code[d] != 1) {
      set_error("binned path deposits take bounded or periodic axes (clamped axes: mpf_deposit_paths)");
      return MPF_ERR_UNSUPPORTED;
    }
  Grid g;
  const int deg = kind == 0 ? 1 : degree;
  int rc = make_grid(g, ndim, kind == 0 ? LINEAR : SPLINE, deg, lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  const int64_t count = g.extent[0] * g.extent[1] * (ndim > 2 ? g.extent[2] : 1) * ndim;
  rc = ensure(ctx->accum, (size_t)count * 8, true, stream);
  if (rc) return rc;
  rc = ensure(ctx->scalars, 64, true, stream);
  if (rc) return rc;
  unsigned long long *maxbits = (unsigned long long *)ctx->scalars.ptr;
  rc = zero_async(maxbits, 8, stream);
  if (rc) return rc;
  const int threads = 256, blocks = blocks_for(n, threads, 4096);
  if (dtype == F32)
    hipLaunchKernelGGL(path_bound_kernel<float>, dim3(blocks), dim3(threads), 0, stream, (const float *)sorted_start,
                       n, (const float *)end, end_stride, (const float *)charges, n, g, maxbits);
  else
    hipLaunchKernelGGL(path_bound_kernel<double>, dim3(blocks), dim3(threads), 0, stream,
                       (const double *)sorted_start, n, (const double *)end, end_stride, (const double *)charges, n,
                       g, maxbits);
  MPF_LAUNCH_CHECK();
  rc = prepare_scale(ctx, nullptr, 0, dtype, 1, n, 1.0, MPF_SCALE_AUTO, stream);
  if (rc) return rc;
  unsigned long long *accum = (unsigned long long *)ctx->accum.ptr;
  const double *slot = scale_slot(ctx);
  const int tdeg = kind == 0 ? 0 : deg;
  if (ndim == 2 && dtype == F32)
    launch_paths_tiled<float, 2>(tdeg, sorted_start, charges, end, end_stride, n, start, g, accum, slot, stream);
  if (ndim == 2 && dtype == F64)
    launch_paths_tiled<double, 2>(tdeg, sorted_start, charges, end, end_stride, n, start, g, accum, slot, stream);
  if (ndim == 3 && dtype == F32)
    launch_paths_tiled<float, 3>(tdeg, sorted_start, charges, end, end_stride, n, start, g, accum, slot, stream);
  if (ndim == 3 && dtype == F64)
    launch_paths_tiled<double, 3>(tdeg, sorted_start, charges, end, end_stride, n, start, g, accum, slot, stream);
  MPF_LAUNCH_CHECK();
  return finalize_fixed(ctx, edges, count, stream);
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

/* The edge field at the markers (the transpose of the path deposits): kind 0 WHITNEY,
 * kind 1 spline edges of `degree`; result (ndim, n) in the marker dtype, overwritten. */
extern "C" int mpf_gather_edges(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, int dtype,
                                int64_t n, const double *edges, void *result, int64_t result_stride,
                                const double *lower, const double *cell_size, const int64_t *num_nodes,
                                const int64_t *code, mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  if (kind != 0 && kind != 1) {
    set_error("edge kind must be 0 (Whitney) or 1 (spline edges), got %d", kind);
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
  if (n < 0) {
    set_error("negative marker count");
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  const int deg = kind == 0 ? 1 : degree;
  int rc = make_grid(g, ndim, kind == 0 ? LINEAR : SPLINE, deg, lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  const int d = kind == 0 ? 0 : degree;
#define GE_ND(T)                                                                                                  \
  if (ndim == 1) launch_gather_edges<T, 1>(d, positions, pos_stride, n, g, edges, result, result_stride, stream); \
  if (ndim == 2) launch_gather_edges<T, 2>(d, positions, pos_stride, n, g, edges, result, result_stride, stream); \
  if (ndim == 3) launch_gather_edges<T, 3>(d, positions, pos_stride, n, g, edges, result, result_stride, stream);
  if (dtype == F32) {
    GE_ND(float)
  } else {
    GE_ND(double)
  }
#undef GE_ND
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

#define GATHER_EDGES_ENTRY(D)                                                                                       \
  extern "C" int mpf_gather_edges_##D##d(MPF_EDGES_GATHER_IN, GEOM##D, MPF_TAIL) {                                  \
    P##D return mpf_gather_edges(0, 1, D, positions, pos_stride, dtype, n, edges, result, result_stride, lo, cs, nn, \
                                 pc, ctx, stream);                                                                \
  }                                                                                                                \
  extern "C" int mpf_gather_edges_spline_##D##d(MPF_EDGES_GATHER_IN, GEOM##D, int degree, MPF_TAIL) {               \
    P##D return mpf_gather_edges(1, degree, D, positions, pos_stride, dtype, n, edges, result, result_stride, lo, cs, \
                                 nn, pc, ctx, stream);                                                            \
  }

GATHER_EDGES_ENTRY(1)
GATHER_EDGES_ENTRY(2)
GATHER_EDGES_ENTRY(3)
