// mpf_forms.hip -- the k-forms of the spline sequence (faces, cells), one work-item per
// marker: scatter_forms_spline_{1,2,3}d_pyccel.py and gather_forms_spline_{1,2,3}d_pyccel.py.
//
// Component c of a k-form is the tensor product of B_p along the axes it is not staggered
// on (spline_axis) and B_{p-1} shifted by half a cell along those it is (along_stencil):
// edges along their axis, faces (k = 2 in 3D) across their normal, cells along every axis
// (spline_edges_pyccel.py:233-245 staggered).  The values are integrals over the k-cells;
// nothing is divided by or multiplied with the measures, so the scatter and the gather
// are each other's transpose.  The position is clamped to [0, num - 1] on a non-periodic
// axis first; a staggered k-cell past a bounded wall does not exist and is skipped (never
// multiplied by zero, so a non-finite field value there stays unread).  The weight
// products run left to right (x, y, z, then the value or the field) and the gather sums in
// the reference's x-outer, z-inner order, so a gathered value is the oracle's to the bit.
#include "mpf_internal.h"

namespace mpf {

template <int NDIM, int FORM>
__host__ __device__ constexpr int form_components() {
  return FORM == NDIM ? 1 : NDIM;
}

template <int NDIM, int FORM>
__host__ __device__ constexpr bool form_staggered(int c, int d) {
  return d >= NDIM ? false : (FORM == NDIM ? true : (FORM == 1 ? d == c : (FORM == 2 && NDIM == 3 ? d != c : false)));
}

// The per-axis stencils of one marker: the spline nodes (DEG + 1, all present) where some
// component is not staggered, the along edges (DEG, some absent) where some component is.
// Axes past NDIM hold one node of weight 1 at index 0 (a multiply by 1.0 is exact).
template <int NDIM, int FORM, int DEG>
struct FormStencil {
  int64_t node[3][DEG + 1], edge[3][DEG];
  double wn[3][DEG + 1], we[3][DEG];
  int ex[3][DEG];

  __device__ __forceinline__ void make(const double *rel, const Grid &g) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      bool need_node = d >= NDIM, need_edge = false;
#pragma unroll
      for (int c = 0; c < form_components<NDIM, FORM>(); ++c) {
        if (form_staggered<NDIM, FORM>(c, d)) need_edge = true;
        else if (d < NDIM) need_node = true;
      }
      if (d >= NDIM) {
        node[d][0] = 0;
        wn[d][0] = 1.0;
        continue;
      }
      if (need_node) spline_axis(rel[d], g.num[d], g.code[d], DEG, node[d], wn[d]);
      if (need_edge) along_stencil(rel[d], g.num[d], g.code[d], DEG, edge[d], we[d], ex[d]);
    }
  }
};

template <int NDIM, int FORM, int DEG, int C, int D>
struct FormAxis {
  static constexpr bool stag = form_staggered<NDIM, FORM>(C, D);
  static constexpr int size = D >= NDIM ? 1 : (stag ? DEG : DEG + 1);
  static __device__ __forceinline__ bool present(const FormStencil<NDIM, FORM, DEG> &s, int a) {
    if constexpr (stag) return s.ex[D][a] != 0;
    else return true;
  }
  static __device__ __forceinline__ double weight(const FormStencil<NDIM, FORM, DEG> &s, int a) {
    if constexpr (stag) return s.we[D][a];
    else return s.wn[D][a];
  }
  static __device__ __forceinline__ int64_t index(const FormStencil<NDIM, FORM, DEG> &s, int a) {
    if constexpr (stag) return s.edge[D][a];
    else return s.node[D][a];
  }
};

// Calls f(i0, i1, i2, weight) for every present k-cell of component C (its index on each
// axis, 0 past NDIM): the weights multiplied x, y, z left to right.
template <int NDIM, int FORM, int DEG, int C, class F>
__device__ __forceinline__ void for_each_cell(const FormStencil<NDIM, FORM, DEG> &s, F &&f) {
  using X = FormAxis<NDIM, FORM, DEG, C, 0>;
  using Y = FormAxis<NDIM, FORM, DEG, C, 1>;
  using Z = FormAxis<NDIM, FORM, DEG, C, 2>;
#pragma unroll
  for (int a = 0; a < X::size; ++a) {
    if (!X::present(s, a)) continue;
    const double wa = X::weight(s, a);
#pragma unroll
    for (int b = 0; b < Y::size; ++b) {
      if (!Y::present(s, b)) continue;
      const double wb = NDIM >= 2 ? wa * Y::weight(s, b) : wa;
#pragma unroll
      for (int e = 0; e < Z::size; ++e) {
        if (!Z::present(s, e)) continue;
        const double w = NDIM == 3 ? wb * Z::weight(s, e) : wb;
        f(X::index(s, a), Y::index(s, b), Z::index(s, e), w);
      }
    }
  }
}

// The position in cells, clamped to [0, num - 1] on a non-periodic axis
// (gather_forms_spline_2d_pyccel.py:52-63; an infinite position lands on the wall).  A NaN,
// or a non-finite position on a periodic axis, for which the reference's integer cast is
// undefined, is taken as 0 (the cell the binning gives it), so that every index stays
// inside the field.
template <int NDIM>
__device__ __forceinline__ void form_relative(const double *x, const Grid &g, double *rel) {
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    double r = (x[d] - g.lower[d]) * g.inverse[d];
    if (g.code[d] == 1) {
      if (!isfinite(r)) r = 0.0;
    } else {
      if (r != r) r = 0.0;
      if (r < 0.0) r = 0.0;
      if (r > (double)(g.num[d] - 1)) r = (double)g.num[d] - 1.0;
    }
    rel[d] = r;
  }
}

template <int NDIM>
__device__ __forceinline__ void form_strides(const Grid &g, int64_t &s0, int64_t &s1) {
  s0 = NDIM == 3 ? g.extent[1] * g.extent[2] : (NDIM == 2 ? g.extent[1] : 1);
  s1 = NDIM == 3 ? g.extent[2] : 1;
}

template <class T, int NDIM, int FORM, int DEG, int C>
__device__ __forceinline__ void gather_component(const FormStencil<NDIM, FORM, DEG> &s, int64_t s0, int64_t s1,
                                                 const double *__restrict__ field, T *__restrict__ res, int64_t rs,
                                                 int64_t m) {
  constexpr int NC = form_components<NDIM, FORM>();
  if constexpr (C < NC) {
    double total = 0.0;
    for_each_cell<NDIM, FORM, DEG, C>(s, [&](int64_t i0, int64_t i1, int64_t i2, double w) {
      total += w * field[(i0 * s0 + i1 * s1 + i2) * NC + C];
    });
    res[C * rs + m] = (T)total;
  }
}

template <class T, int NDIM, int FORM, int DEG>
__global__ void __launch_bounds__(256) gather_forms_kernel(const T *__restrict__ pos, int64_t ps, int64_t n, Grid g,
                                                           const double *__restrict__ field, T *__restrict__ res,
                                                           int64_t rs) {
  int64_t s0, s1;
  form_strides<NDIM>(g, s0, s1);
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    double x[NDIM], rel[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) x[d] = (double)pos[d * ps + m];
    form_relative<NDIM>(x, g, rel);
    FormStencil<NDIM, FORM, DEG> s;
    s.make(rel, g);
    gather_component<T, NDIM, FORM, DEG, 0>(s, s0, s1, field, res, rs, m);
    gather_component<T, NDIM, FORM, DEG, 1>(s, s0, s1, field, res, rs, m);
    gather_component<T, NDIM, FORM, DEG, 2>(s, s0, s1, field, res, rs, m);
  }
}

template <int MODE, int NDIM, int FORM, int DEG, int C>
__device__ __forceinline__ void scatter_component(const FormStencil<NDIM, FORM, DEG> &s, int64_t s0, int64_t s1,
                                                  double value, double scale, double *__restrict__ field,
                                                  unsigned long long *__restrict__ accum) {
  constexpr int NC = form_components<NDIM, FORM>();
  if constexpr (C < NC) {
    for_each_cell<NDIM, FORM, DEG, C>(s, [&](int64_t i0, int64_t i1, int64_t i2, double w) {
      const int64_t at = i0 * s0 + i1 * s1 + i2;
      const double contribution = w * value;
      if (MODE == MODE_ATOMIC)
        unsafeAtomicAdd(field + at * NC + C, contribution);
      else
        atomicAdd(accum + at * NC + C, to_fixed(contribution * scale));
    });
  }
}

template <class T, int NDIM, int FORM, int DEG, int MODE>
__global__ void __launch_bounds__(256) scatter_forms_kernel(const T *__restrict__ pos, int64_t ps,
                                                            const T *__restrict__ val, int64_t vs, int64_t n, Grid g,
                                                            double *__restrict__ field,
                                                            unsigned long long *__restrict__ accum,
                                                            const double *__restrict__ scale_slot) {
  constexpr int NC = form_components<NDIM, FORM>();
  const double scale = MODE == MODE_FIXED ? scale_slot[0] : 1.0;
  int64_t s0, s1;
  form_strides<NDIM>(g, s0, s1);
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    double v[NC];
    bool live = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      v[c] = (double)val[c * vs + m];
      live |= v[c] != 0.0;
    }
    if (!live) continue;  // a dead marker: its position is not read
    double x[NDIM], rel[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) x[d] = (double)pos[d * ps + m];
    form_relative<NDIM>(x, g, rel);
    FormStencil<NDIM, FORM, DEG> s;
    s.make(rel, g);
    scatter_component<MODE, NDIM, FORM, DEG, 0>(s, s0, s1, v[0], scale, field, accum);
    if constexpr (NC > 1) scatter_component<MODE, NDIM, FORM, DEG, 1>(s, s0, s1, v[NC > 1 ? 1 : 0], scale, field, accum);
    if constexpr (NC > 2) scatter_component<MODE, NDIM, FORM, DEG, 2>(s, s0, s1, v[NC > 2 ? 2 : 0], scale, field, accum);
  }
}

// ---- the tiled deposit of binned markers --------------------------------------------------
//
// Markers binned by the anchor cell of the degree-DEG spline nodes (mpf_bin_markers, kind
// SPLINE), the same tiles as the node kernels: one workgroup per tile, an int64
// fixed-point accumulator over the tile's footprint in LDS, each work-item a contiguous run
// of the tile's markers adding to_fixed(w * v * 2^k) per k-cell -- the integers of the
// one-marker-per-work-item fixed-point deposit, so the field is the same bit for bit --
// then the tile's nonzero words into the global int64 accumulator.  Along each axis a
// marker's nodes are anchor - DEG/2 .. anchor - DEG/2 + DEG and its staggered k-cells lie
// among them, so the footprint is the tile plus DEG/2 below and DEG - DEG/2 above; the one
// exception, an edge stencil clamped at a bounded wall that reaches one node lower, leaves
// the footprint only in a tile that starts at the wall's last node, and goes to the global
// accumulator.  A periodic axis folds the footprint (it needs num >= the footprint); a
// bounded one clamps.  The tiles add their words with integer atomics (exact, so the field
// does not depend on the order); running overlapping tiles in separate launches without
// atomics measured slower (the launches' tails cost more than the atomics).
template <int NDIM, int DEG>
struct FormTile {
  static constexpr int TX = NDIM == 3 ? TILE_X : TILE2, TY = NDIM == 3 ? TILE_Y : TILE2, TZ = NDIM == 3 ? TILE_Z : 1;
  static constexpr int TC = TX * TY * TZ;
  static constexpr int LO = DEG / 2;                 // footprint nodes below the tile
  static constexpr int SPAN = DEG;                   // footprint nodes beyond the tile's own
  static constexpr int LX = TX + SPAN, LY = TY + SPAN, LZ = NDIM == 3 ? TZ + SPAN : 1;
};

// The footprint slot of node k (already wrapped or clamped) of an axis whose footprint
// starts at unwrapped node `base`; -1 when outside (which the stencils never give).
__device__ __forceinline__ int form_slot(int64_t k, int base, int len, int64_t num, int code) {
  int64_t s = k - base;
  if (code == 1) s = s < 0 ? s + num : (s >= num ? s - num : s);
  return (s >= 0 && s < len) ? (int)s : -1;
}

// Component C's present k-cells of one marker as tile slots: per axis the slot offset of
// each stencil entry (x: slot * LY * LZ, y: slot * LZ, z: slot), its weight, and whether it
// exists; `inside` is false when some entry falls outside the footprint.
template <int NDIM, int FORM, int DEG, int C, int LX, int LY, int LZ>
struct TileCells {
  using X = FormAxis<NDIM, FORM, DEG, C, 0>;
  using Y = FormAxis<NDIM, FORM, DEG, C, 1>;
  using Z = FormAxis<NDIM, FORM, DEG, C, 2>;
  int ox[X::size], oy[Y::size], oz[Z::size];
  bool inside = true;

  __device__ __forceinline__ TileCells(const FormStencil<NDIM, FORM, DEG> &s, const Grid &g, int bx, int by, int bz) {
#pragma unroll
    for (int a = 0; a < X::size; ++a) {
      const int sl = form_slot(X::index(s, a), bx, LX, g.num[0], g.code[0]);
      inside = inside && (sl >= 0 || !X::present(s, a));
      ox[a] = sl * (LY * LZ);
    }
#pragma unroll
    for (int b = 0; b < Y::size; ++b) {
      const int sl = NDIM >= 2 ? form_slot(Y::index(s, b), by, LY, g.num[1], g.code[1]) : 0;
      inside = inside && (sl >= 0 || !Y::present(s, b));
      oy[b] = sl * LZ;
    }
#pragma unroll
    for (int e = 0; e < Z::size; ++e) {
      const int sl = NDIM == 3 ? form_slot(Z::index(s, e), bz, LZ, g.num[2], g.code[2]) : 0;
      inside = inside && (sl >= 0 || !Z::present(s, e));
      oz[e] = sl;
    }
  }
};

template <class T, int NDIM, int FORM, int DEG, int BLOCK>
__global__ void __launch_bounds__(BLOCK) scatter_forms_tiled(const T *__restrict__ pos, const T *__restrict__ val,
                                                             int64_t n, const uint32_t *__restrict__ start, Grid g,
                                                             int nty, int ntz, unsigned long long *__restrict__ accum,
                                                             const double *__restrict__ scalars) {
  using Ti = FormTile<NDIM, DEG>;
  constexpr int NC = form_components<NDIM, FORM>();
  constexpr int LX = Ti::LX, LY = Ti::LY, LZ = Ti::LZ, VOL = LX * LY * LZ, WORDS = VOL * NC;
  __shared__ unsigned long long acc[WORDS];  // component-major: acc[c * VOL + slot]
  const int tile = blockIdx.x, t = threadIdx.x;
  const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
  const int bx = tx * Ti::TX - Ti::LO, by = ty * Ti::TY - Ti::LO, bz = NDIM == 3 ? tz * Ti::TZ - Ti::LO : 0;
  for (int i = t; i < WORDS; i += BLOCK) acc[i] = 0ull;
  const uint32_t first = start[(int64_t)tile * Ti::TC], last = start[(int64_t)(tile + 1) * Ti::TC];
  const double scale = scalars[0];
  int64_t s0, s1;
  form_strides<NDIM>(g, s0, s1);
  __syncthreads();
  // 3D tiles: an even split of the tile's few thousand markers, one run per lane; 2D tiles
  // (tens of thousands): runs of 8 consecutive markers, run t of each pass of BLOCK runs, so
  // a wavefront's row loads stay coalesced
  const uint32_t count = last - first;
  const uint32_t run = NDIM == 2 ? 8u : (count + BLOCK - 1) / BLOCK;
  for (uint32_t lo = first + t * run; lo < last; lo += BLOCK * run)
  for (uint32_t m = lo; m < (lo + run < last ? lo + run : last); ++m) {
    double x[NDIM], rel[NDIM], vs[NC];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) x[d] = (double)pos[d * n + m];
    // value * 2^k: a power-of-two scale, so round(w * value) * 2^k == round(w * (value * 2^k))
    // and the integers are the direct deposit's
    double vmax = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      vs[c] = (double)val[c * n + m] * scale;
      vmax = fmax(vmax, fabs(vs[c]));
    }
    form_relative<NDIM>(x, g, rel);
    FormStencil<NDIM, FORM, DEG> s;
    s.make(rel, g);
    // the weights are at most 1 off clamped axes: below 2^50 quanta the contributions take
    // the two-operation rounding (the scale the binning picks keeps a value below 2^49)
    const bool small = vmax < 0x1p50;
    auto add = [&](auto cc, auto fast) {
      constexpr int C = decltype(cc)::value;
      if constexpr (C < NC) {
        const double v = vs[C];
        TileCells<NDIM, FORM, DEG, C, LX, LY, LZ> cells(s, g, bx, by, bz);
        using X = typename decltype(cells)::X;
        using Y = typename decltype(cells)::Y;
        using Z = typename decltype(cells)::Z;
        if (cells.inside) {
#pragma unroll
          for (int a = 0; a < X::size; ++a) {
            if (!X::present(s, a)) continue;
            const double wa = X::weight(s, a);
#pragma unroll
            for (int b = 0; b < Y::size; ++b) {
              if (!Y::present(s, b)) continue;
              const double wb = NDIM >= 2 ? wa * Y::weight(s, b) : wa;
              const int oab = C * VOL + cells.ox[a] + cells.oy[b];
#pragma unroll
              for (int e = 0; e < Z::size; ++e) {
                if (!Z::present(s, e)) continue;
                const double y = (NDIM == 3 ? wb * Z::weight(s, e) : wb) * v;
                atomicAdd(acc + oab + cells.oz[e], decltype(fast)::value ? to_fixed_small(y) : to_fixed(y));
              }
            }
          }
        } else {  // outside the footprint (which the stencils never give): the global accumulator
          for_each_cell<NDIM, FORM, DEG, C>(s, [&](int64_t i0, int64_t i1, int64_t i2, double w) {
            atomicAdd(accum + (i0 * s0 + i1 * s1 + i2) * NC + C, to_fixed(w * v));
          });
        }
      }
    };
    auto add_all = [&](auto fast) {
      add(std::integral_constant<int, 0>{}, fast);
      add(std::integral_constant<int, 1>{}, fast);
      add(std::integral_constant<int, 2>{}, fast);
    };
    if (small)
      add_all(std::true_type{});
    else
      add_all(std::false_type{});
  }
  __syncthreads();
  for (int i = t; i < WORDS; i += BLOCK) {
    const unsigned long long q = acc[i];
    if (q == 0ull) continue;
    const int c = i / VOL, sl = i - c * VOL, sz = sl % LZ, sy = (sl / LZ) % LY, sx = sl / (LZ * LY);
    int64_t gx = bx + sx, gy = by + sy, gz = bz + sz;
    if (g.code[0] == 1) gx = gx < 0 ? gx + g.num[0] : (gx >= g.num[0] ? gx - g.num[0] : gx);
    if (NDIM >= 2 && g.code[1] == 1) gy = gy < 0 ? gy + g.num[1] : (gy >= g.num[1] ? gy - g.num[1] : gy);
    if (NDIM == 3 && g.code[2] == 1) gz = gz < 0 ? gz + g.num[2] : (gz >= g.num[2] ? gz - g.num[2] : gz);
    atomicAdd(accum + (gx * s0 + gy * s1 + gz) * NC + c, q);
  }
}

// The gather of binned markers: one workgroup per tile stages the tile's footprint of the
// field in LDS (component-major; slots past a bounded wall hold 0 and are never read), then
// each lane gathers one marker at a time (consecutive lanes on consecutive markers), with
// the direct gather's products and sums in its order -- the same values bit for bit.
// Results in bins order.
template <class T, int NDIM, int FORM, int DEG, int BLOCK>
__global__ void __launch_bounds__(BLOCK) gather_forms_tiled(const T *__restrict__ pos, int64_t n,
                                                            const uint32_t *__restrict__ start, Grid g, int nty,
                                                            int ntz, const double *__restrict__ field,
                                                            T *__restrict__ res) {
  using Ti = FormTile<NDIM, DEG>;
  constexpr int NC = form_components<NDIM, FORM>();
  constexpr int LX = Ti::LX, LY = Ti::LY, LZ = Ti::LZ, VOL = LX * LY * LZ;
  __shared__ double tile_f[VOL * NC];
  const int tile = blockIdx.x, t = threadIdx.x;
  const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
  const int bx = tx * Ti::TX - Ti::LO, by = ty * Ti::TY - Ti::LO, bz = NDIM == 3 ? tz * Ti::TZ - Ti::LO : 0;
  int64_t s0, s1;
  form_strides<NDIM>(g, s0, s1);
  for (int i = t; i < VOL; i += BLOCK) {
    const int sz = i % LZ, sy = (i / LZ) % LY, sx = i / (LZ * LY);
    int64_t k[3] = {bx + sx, by + sy, bz + sz};
    bool ok = true;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      if (g.code[d] == 1) k[d] = k[d] < 0 ? k[d] + g.num[d] : (k[d] >= g.num[d] ? k[d] - g.num[d] : k[d]);
      ok = ok && k[d] >= 0 && k[d] < g.extent[d];
    }
    const int64_t at = (k[0] * s0 + k[1] * s1 + k[2]) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) tile_f[c * VOL + i] = ok ? field[at + c] : 0.0;
  }
  const uint32_t first = start[(int64_t)tile * Ti::TC], last = start[(int64_t)(tile + 1) * Ti::TC];
  __syncthreads();
  for (uint32_t m = first + t; m < last; m += BLOCK) {
    double x[NDIM], rel[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) x[d] = (double)pos[d * n + m];
    form_relative<NDIM>(x, g, rel);
    FormStencil<NDIM, FORM, DEG> s;
    s.make(rel, g);
    auto one = [&](auto cc) {
      constexpr int C = decltype(cc)::value;
      if constexpr (C < NC) {
        TileCells<NDIM, FORM, DEG, C, LX, LY, LZ> cells(s, g, bx, by, bz);
        using X = typename decltype(cells)::X;
        using Y = typename decltype(cells)::Y;
        using Z = typename decltype(cells)::Z;
        double total = 0.0;
        if (cells.inside) {
#pragma unroll
          for (int a = 0; a < X::size; ++a) {
            if (!X::present(s, a)) continue;
            const double wa = X::weight(s, a);
#pragma unroll
            for (int b = 0; b < Y::size; ++b) {
              if (!Y::present(s, b)) continue;
              const double wb = NDIM >= 2 ? wa * Y::weight(s, b) : wa;
              const int oab = C * VOL + cells.ox[a] + cells.oy[b];
#pragma unroll
              for (int e = 0; e < Z::size; ++e) {
                if (!Z::present(s, e)) continue;
                const double w = NDIM == 3 ? wb * Z::weight(s, e) : wb;
                total += w * tile_f[oab + cells.oz[e]];
              }
            }
          }
        } else {  // outside the footprint (which the stencils never give): the global field
          for_each_cell<NDIM, FORM, DEG, C>(s, [&](int64_t i0, int64_t i1, int64_t i2, double w) {
            total += w * field[(i0 * s0 + i1 * s1 + i2) * NC + C];
          });
        }
        res[(int64_t)C * n + m] = (T)total;
      }
    };
    one(std::integral_constant<int, 0>{});
    one(std::integral_constant<int, 1>{});
    one(std::integral_constant<int, 2>{});
  }
}

// The markers of the trailing dead bin (binned with all-zero values): the direct gather.
template <class T, int NDIM, int FORM, int DEG>
__global__ void __launch_bounds__(256) gather_forms_dead(const T *__restrict__ pos, int64_t n,
                                                         const uint32_t *__restrict__ dead_first, Grid g,
                                                         const double *__restrict__ field, T *__restrict__ res) {
  int64_t s0, s1;
  form_strides<NDIM>(g, s0, s1);
  for (int64_t m = (int64_t)dead_first[0] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n;
       m += (int64_t)gridDim.x * blockDim.x) {
    double x[NDIM], rel[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) x[d] = (double)pos[d * n + m];
    form_relative<NDIM>(x, g, rel);
    FormStencil<NDIM, FORM, DEG> s;
    s.make(rel, g);
    gather_component<T, NDIM, FORM, DEG, 0>(s, s0, s1, field, res, n, m);
    gather_component<T, NDIM, FORM, DEG, 1>(s, s0, s1, field, res, n, m);
    gather_component<T, NDIM, FORM, DEG, 2>(s, s0, s1, field, res, n, m);
  }
}

// 512 lanes per tile: the 2 tiles a CU's LDS holds then keep 16 waves in flight (measured:
// 256 lanes 2.0 ms, 512 1.2 ms, 1024 1.4 ms for 2^25 markers, 3D faces of degree 2, 256^3)
constexpr int GATHER_BLOCK = 512;
template <class T, int NDIM>
static int launch_gather_forms_tiled(int form, int degree, const void *pos, int64_t n, const uint32_t *start,
                                     const Grid &g, const double *field, void *res, hipStream_t stream) {
  const int tx = NDIM == 3 ? TILE_X : TILE2, ty = NDIM == 3 ? TILE_Y : TILE2;
  const int ntx = (int)((g.num[0] + tx - 1) / tx), nty = (int)((g.num[1] + ty - 1) / ty);
  const int ntz = NDIM == 3 ? (int)((g.num[2] + TILE_Z - 1) / TILE_Z) : 1;
  const int blocks = ntx * nty * ntz;
  const uint32_t *dead = start + (int64_t)blocks * (NDIM == 3 ? TILE_CELLS : TILE2_CELLS);
  const int dead_blocks = blocks_for(n, 256, 1024);
#define GT(F, D)                                                                                                  \
  hipLaunchKernelGGL((gather_forms_tiled<T, NDIM, F, D, GATHER_BLOCK>), dim3(blocks), dim3(GATHER_BLOCK), 0, stream, \
                     (const T *)pos, n,                                                                            \
                     start, g, nty, ntz, field, (T *)res);                                                         \
  hipLaunchKernelGGL((gather_forms_dead<T, NDIM, F, D>), dim3(dead_blocks), dim3(256), 0, stream, (const T *)pos,  \
                     n, dead, g, field, (T *)res)
#define GT_DEG(F)       \
  if (degree == 1) {    \
    GT(F, 1);           \
  } else if (degree == 2) { \
    GT(F, 2);           \
  } else {              \
    GT(F, 3);           \
  }
  if (form == 1) {
    GT_DEG(1)
  } else if (form == 2) {
    GT_DEG(2)
  } else if constexpr (NDIM == 3) {
    GT_DEG(3)
  }
#undef GT_DEG
#undef GT
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

// Whether the tiled deposit takes this case: 2D or 3D, no clamped axis, and a periodic
// axis at least as long as the footprint.
static bool forms_tiled_supported(int degree, int ndim, const Grid &g) {
  if (ndim != 2 && ndim != 3) return false;
  const int span = degree;
  const int tile[3] = {ndim == 3 ? TILE_X : TILE2, ndim == 3 ? TILE_Y : TILE2, TILE_Z};
  for (int d = 0; d < ndim; ++d) {
    if (g.code[d] == 2) return false;
    if (g.code[d] == 1 && g.num[d] < tile[d] + span) return false;
  }
  return true;
}

// 512 lanes per tile (the LDS holds 2 tiles per CU; measured for 2^25 markers, 3D faces of
// degree 2, 256^3: 256 lanes 2.4 ms, 512 1.8 ms, 1024 2.4 ms)
constexpr int SCATTER_BLOCK = 512;
template <class T, int NDIM>
static int launch_forms_tiled(int form, int degree, const void *pos, const void *val, int64_t n,
                              const uint32_t *start, const Grid &g, unsigned long long *accum,
                              const double *scalars, hipStream_t stream) {
  const int tx = NDIM == 3 ? TILE_X : TILE2, ty = NDIM == 3 ? TILE_Y : TILE2;
  const int ntx = (int)((g.num[0] + tx - 1) / tx), nty = (int)((g.num[1] + ty - 1) / ty);
  const int ntz = NDIM == 3 ? (int)((g.num[2] + TILE_Z - 1) / TILE_Z) : 1;
  const int blocks = ntx * nty * ntz;
#define FT(F, D)                                                                                                  \
  hipLaunchKernelGGL((scatter_forms_tiled<T, NDIM, F, D, SCATTER_BLOCK>), dim3(blocks), dim3(SCATTER_BLOCK), 0,  \
                     stream, (const T *)pos,                                                                       \
                     (const T *)val, n, start, g, nty, ntz, accum, scalars)
#define FT_DEG(F)       \
  if (degree == 1)      \
    FT(F, 1);           \
  else if (degree == 2) \
    FT(F, 2);           \
  else                  \
    FT(F, 3);
  if (form == 1) {
    FT_DEG(1)
  } else if (form == 2) {
    FT_DEG(2)
  } else if constexpr (NDIM == 3) {
    FT_DEG(3)
  }
#undef FT_DEG
#undef FT
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

template <class T, int NDIM, int FORM>
static void launch_forms(bool gather, int mode, int degree, const void *pos, int64_t ps, const void *val, int64_t vs,
                         int64_t n, const Grid &g, const double *cfield, double *field, void *res, int64_t rs,
                         unsigned long long *accum, const double *slot, hipStream_t stream) {
  const int threads = 256, blocks = blocks_for(n, threads, 1 << 16);
#define FORM_LAUNCH(DEG)                                                                                           \
  if (gather)                                                                                                      \
    hipLaunchKernelGGL((gather_forms_kernel<T, NDIM, FORM, DEG>), dim3(blocks), dim3(threads), 0, stream,          \
                       (const T *)pos, ps, n, g, cfield, (T *)res, rs);                                            \
  else if (mode == MODE_ATOMIC)                                                                                    \
    hipLaunchKernelGGL((scatter_forms_kernel<T, NDIM, FORM, DEG, MODE_ATOMIC>), dim3(blocks), dim3(threads), 0,    \
                       stream, (const T *)pos, ps, (const T *)val, vs, n, g, field, accum, slot);                  \
  else                                                                                                             \
    hipLaunchKernelGGL((scatter_forms_kernel<T, NDIM, FORM, DEG, MODE_FIXED>), dim3(blocks), dim3(threads), 0,     \
                       stream, (const T *)pos, ps, (const T *)val, vs, n, g, field, accum, slot);
  if (degree == 1) {
    FORM_LAUNCH(1)
  } else if (degree == 2) {
    FORM_LAUNCH(2)
  } else {
    FORM_LAUNCH(3)
  }
#undef FORM_LAUNCH
}

template <class T>
static void dispatch_forms(bool gather, int mode, int degree, int form, int ndim, const void *pos, int64_t ps,
                           const void *val, int64_t vs, int64_t n, const Grid &g, const double *cfield, double *field,
                           void *res, int64_t rs, unsigned long long *accum, const double *slot, hipStream_t stream) {
#define FORM_ARGS gather, mode, degree, pos, ps, val, vs, n, g, cfield, field, res, rs, accum, slot, stream
  if (ndim == 1) launch_forms<T, 1, 1>(FORM_ARGS);
  if (ndim == 2 && form == 1) launch_forms<T, 2, 1>(FORM_ARGS);
  if (ndim == 2 && form == 2) launch_forms<T, 2, 2>(FORM_ARGS);
  if (ndim == 3 && form == 1) launch_forms<T, 3, 1>(FORM_ARGS);
  if (ndim == 3 && form == 2) launch_forms<T, 3, 2>(FORM_ARGS);
  if (ndim == 3 && form == 3) launch_forms<T, 3, 3>(FORM_ARGS);
#undef FORM_ARGS
}

// spline_forms.py:46-55 _check_form and spline_edges.py:31-34 _check.
static int check_forms(int degree, int form, int ndim, int dtype, int64_t n, int64_t ncomp) {
  if (degree < 1 || degree > 3) {
    set_error("spline edge degrees are (1, 2, 3), got %d", degree);
    return MPF_ERR_ARGUMENT;
  }
  if (ndim < 1 || ndim > 3) {
    set_error("ndim must be 1, 2 or 3, got %d", ndim);
    return MPF_ERR_ARGUMENT;
  }
  if (form < 1 || form > ndim) {
    set_error("forms of %d dimensions are 1 to %d, got %d", ndim, ndim, form);
    return MPF_ERR_ARGUMENT;
  }
  const int64_t want = form == ndim ? 1 : ndim;
  if (ncomp != want) {
    set_error("a %d-form in %dD has %lld components, got %lld", form, ndim, (long long)want, (long long)ncomp);
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
  return MPF_OK;
}

}  // namespace mpf

using namespace mpf;

extern "C" int mpf_scatter_forms(int degree, int form, int ndim, const void *positions, int64_t pos_stride,
                                 const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp,
                                 double *field, const double *lower, const double *cell_size,
                                 const int64_t *num_nodes, const int64_t *code, int mode, int scale_exp,
                                 mpf_context *ctx, mpf_stream stream_) {
  int rc = check_forms(degree, form, ndim, dtype, n, ncomp);
  if (rc) return rc;
  if (mode != MODE_ATOMIC && mode != MODE_FIXED) {
    set_error("mode must be MPF_MODE_ATOMIC or MPF_MODE_FIXED, got %d", mode);
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  rc = make_grid(g, ndim, SPLINE, degree, lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  const int64_t count = g.extent[0] * (ndim > 1 ? g.extent[1] : 1) * (ndim > 2 ? g.extent[2] : 1) * ncomp;
  unsigned long long *accum = nullptr;
  const double *slot = nullptr;
  if (mode == MODE_FIXED) {
    rc = ensure(ctx->accum, (size_t)count * 8, true, stream);
    if (rc) return rc;
    // a staggered weight on a clamped axis is a B-spline rescaled by degree / (knot span),
    // at most `degree`; elsewhere the weights of a marker sum to at most one
    double bound = 1.0;
    for (int d = 0; d < ndim; ++d)
      if (code[d] == 2) bound *= (double)degree;
    rc = prepare_scale(ctx, values, val_stride, dtype, ncomp, n, bound, scale_exp, stream);
    if (rc) return rc;
    accum = (unsigned long long *)ctx->accum.ptr;
    slot = scale_slot(ctx);
  }
  {
    ProfileScope prof("scatter_forms", stream);
    if (dtype == F32)
      dispatch_forms<float>(false, mode, degree, form, ndim, positions, pos_stride, values, val_stride, n, g, nullptr,
                            field, nullptr, 0, accum, slot, stream);
    else
      dispatch_forms<double>(false, mode, degree, form, ndim, positions, pos_stride, values, val_stride, n, g,
                             nullptr, field, nullptr, 0, accum, slot, stream);
  }
  MPF_LAUNCH_CHECK();
  if (mode == MODE_FIXED) return finalize_fixed(ctx, field, count, stream);
  return MPF_OK;
}

extern "C" int mpf_gather_forms(int degree, int form, int ndim, const void *positions, int64_t pos_stride, int dtype,
                                int64_t n, const double *field, int64_t ncomp, void *result, int64_t result_stride,
                                const double *lower, const double *cell_size, const int64_t *num_nodes,
                                const int64_t *code, mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  int rc = check_forms(degree, form, ndim, dtype, n, ncomp);
  if (rc) return rc;
  Grid g;
  rc = make_grid(g, ndim, SPLINE, degree, lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  ProfileScope prof("gather_forms", stream);
  if (dtype == F32)
    dispatch_forms<float>(true, MODE_ATOMIC, degree, form, ndim, positions, pos_stride, nullptr, 0, n, g, field,
                          nullptr, result, result_stride, nullptr, nullptr, stream);
  else
    dispatch_forms<double>(true, MODE_ATOMIC, degree, form, ndim, positions, pos_stride, nullptr, 0, n, g, field,
                           nullptr, result, result_stride, nullptr, nullptr, stream);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

extern "C" int mpf_scatter_forms_binned(int degree, int form, int ndim, const void *sorted_positions,
                                        const void *sorted_values, int dtype, int64_t n, int64_t ncomp,
                                        const void *start, const double *scalars, double *field,
                                        const double *lower, const double *cell_size, const int64_t *num_nodes,
                                        const int64_t *code, mpf_context *ctx, mpf_stream stream_) {
  int rc = check_forms(degree, form, ndim, dtype, n, ncomp);
  if (rc) return rc;
  Grid g;
  rc = make_grid(g, ndim, SPLINE, degree, lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (!forms_tiled_supported(degree, ndim, g)) {
    set_error("the tiled forms deposit takes 2D and 3D grids without clamped axes whose periodic axes hold at "
              "least a tile and %d nodes", degree);
    return MPF_ERR_UNSUPPORTED;
  }
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  const int64_t count = g.extent[0] * g.extent[1] * (ndim > 2 ? g.extent[2] : 1) * ncomp;
  rc = ensure(ctx->accum, (size_t)count * 8, true, stream);
  if (rc) return rc;
  unsigned long long *accum = (unsigned long long *)ctx->accum.ptr;
  const uint32_t *bins = (const uint32_t *)start;
  {
    ProfileScope prof("scatter_forms_tiled", stream);
    if (ndim == 2 && dtype == F32)
      rc = launch_forms_tiled<float, 2>(form, degree, sorted_positions, sorted_values, n, bins, g, accum, scalars,
                                        stream);
    if (ndim == 2 && dtype == F64)
      rc = launch_forms_tiled<double, 2>(form, degree, sorted_positions, sorted_values, n, bins, g, accum, scalars,
                                         stream);
    if (ndim == 3 && dtype == F32)
      rc = launch_forms_tiled<float, 3>(form, degree, sorted_positions, sorted_values, n, bins, g, accum, scalars,
                                        stream);
    if (ndim == 3 && dtype == F64)
      rc = launch_forms_tiled<double, 3>(form, degree, sorted_positions, sorted_values, n, bins, g, accum, scalars,
                                         stream);
  }
  if (rc) return rc;
  return finalize_fixed_with(ctx, field, count, scalars, stream);
}

extern "C" int mpf_gather_forms_binned(int degree, int form, int ndim, const void *sorted_positions, int dtype,
                                       int64_t n, const void *start, const double *field, int64_t ncomp,
                                       void *result, const double *lower, const double *cell_size,
                                       const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                                       mpf_stream stream_) {
  (void)ctx;
  int rc = check_forms(degree, form, ndim, dtype, n, ncomp);
  if (rc) return rc;
  Grid g;
  rc = make_grid(g, ndim, SPLINE, degree, lower, cell_size, num_nodes, code);
  if (rc) return rc;
  if (!forms_tiled_supported(degree, ndim, g)) {
    set_error("the tiled forms gather takes 2D and 3D grids without clamped axes whose periodic axes hold at "
              "least a tile and %d nodes", degree);
    return MPF_ERR_UNSUPPORTED;
  }
  if (n == 0) return MPF_OK;
  hipStream_t stream = (hipStream_t)stream_;
  const uint32_t *bins = (const uint32_t *)start;
  ProfileScope prof("gather_forms_tiled", stream);
  if (ndim == 2 && dtype == F32)
    rc = launch_gather_forms_tiled<float, 2>(form, degree, sorted_positions, n, bins, g, field, result, stream);
  if (ndim == 2 && dtype == F64)
    rc = launch_gather_forms_tiled<double, 2>(form, degree, sorted_positions, n, bins, g, field, result, stream);
  if (ndim == 3 && dtype == F32)
    rc = launch_gather_forms_tiled<float, 3>(form, degree, sorted_positions, n, bins, g, field, result, stream);
  if (ndim == 3 && dtype == F64)
    rc = launch_gather_forms_tiled<double, 3>(form, degree, sorted_positions, n, bins, g, field, result, stream);
  return rc;
}

#define GEOM1 MPF_GEOM_1D
#define GEOM2 MPF_GEOM_2D
#define GEOM3 MPF_GEOM_3D
#define F1 const double lo[1] = {lower_x}; const double cs[1] = {cell_size_x}; const int64_t nn[1] = {num_x}; \
  const int64_t pc[1] = {periodic_x};
#define F2 const double lo[2] = {lower_x, lower_y}; const double cs[2] = {cell_size_x, cell_size_y}; \
  const int64_t nn[2] = {num_x, num_y}; const int64_t pc[2] = {periodic_x, periodic_y};
#define F3 const double lo[3] = {lower_x, lower_y, lower_z}; \
  const double cs[3] = {cell_size_x, cell_size_y, cell_size_z}; const int64_t nn[3] = {num_x, num_y, num_z}; \
  const int64_t pc[3] = {periodic_x, periodic_y, periodic_z};

#define FORMS_ENTRY(D)                                                                                              \
  extern "C" int mpf_scatter_forms_spline_##D##d(MPF_SCATTER_IN, GEOM##D, int degree, int form, int mode,           \
                                                 MPF_TAIL) {                                                      \
    F##D return mpf_scatter_forms(degree, form, D, positions, pos_stride, values, val_stride, dtype, n, ncomp,    \
                                  field, lo, cs, nn, pc, mode, MPF_SCALE_AUTO, ctx, stream);                      \
  }                                                                                                              \
  extern "C" int mpf_gather_forms_spline_##D##d(MPF_GATHER_IN, GEOM##D, int degree, int form, MPF_TAIL) {          \
    F##D return mpf_gather_forms(degree, form, D, positions, pos_stride, dtype, n, field, ncomp, result,            \
                                 result_stride, lo, cs, nn, pc, ctx, stream);                                     \
  }

FORMS_ENTRY(1)
FORMS_ENTRY(2)
FORMS_ENTRY(3)
