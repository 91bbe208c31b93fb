// mpf_binned.hip -- the hot path: markers binned by stencil cell with an on-device
// counting sort, then deposited and gathered one tile of cells per workgroup.
//
// Binning.  A marker's stencil cell ("anchor") is the cell whose nodes it touches, found
// with the same float64 arithmetic as the reference's stencils (quadratic:
// int(r + 1/2); linear: the left node; spline p: int(r + 1/2) for even p, int(r) for odd p;
// wrapped on a periodic axis, clamped on a bounded one).  3D keys put the cells of a tile
// next to each other (tile order), so that a tile's markers are one contiguous run; a
// count pass (atomic ranks), an exclusive scan and a placement pass give the markers
// sorted by key, the bin offsets and the permutation.  Markers whose values are all zero
// (dead) go to a trailing bin that deposits nothing.
//
// Deposit.  One workgroup per tile of anchor cells, in parts of 256 cells whose markers are
// staged in LDS.  The markers of a cell are cut into items of about three; a lane runs over
// an item and accumulates its (p+1)^3 x C contributions in float64 registers as exact
// integers: every contribution is rounded to an integer number of fixed-point quanta
// inside one fma on an accumulator biased by 1.5 * 2^52, where the unit in the last place
// is one quantum; the factors are arranged so that the exact product is never a tie (see
// odd_significand below), hence the sum is independent of the marker order.  Each lane
// then adds its integers into the tile's int64 LDS accumulator (covering the tile plus its
// stencil halo), and the workgroup adds the tile into the float64 field once.  Tiles whose
// footprints overlap run in different launches (a 2- or 3-colouring per axis), in a fixed
// order, so the field update needs no atomics and is bit-reproducible.
//
// Gather.  One workgroup per tile: the tile's node footprint is staged in LDS and each
// lane interpolates one marker at a time, with the weight products in the reference's
// order (gather_quadratic_3d_pyccel.py:103-114).
#include <hipcub/hipcub.hpp>

#include <atomic>
#include <map>
#include <mutex>
#include <tuple>

#include "mpf_internal.h"

namespace mpf {

// A uniform value held in scalar registers from this point on: the wait for its scalar load
// is placed here, once, instead of at its first use inside a loop, where it would be taken
// on every trip and, since scalar loads and LDS operations share one counter, would also
// wait for every LDS operation the loop had issued.
template <class T>
__device__ __forceinline__ T uniform_here(T v) {
  asm volatile("" : "+s"(v));
  return v;
}

// ---- per-axis stencil cell and weights -----------------------------------------------

// The cell of rel on a clamped axis (clamped_cell's: the position clamped to [0, num - 1],
// the cell to [0, num - 2]); its stencil nodes are cell .. cell + P of num + P - 1.
__device__ __forceinline__ int clamped_axis_cell(double rel, int num) {
  double r = rel;
  if (r < 0.0) r = 0.0;
  if (r > (double)(num - 1)) r = (double)num - 1.0;  // NaN stays NaN: the cast below gives 0
  const int cell = (int)fmax(r, 0.0);
  return cell > num - 2 ? num - 2 : cell;
}

// The first stencil node of a tile cell's stencil relative to the cell: -OFF for the
// uniform stencils (anchor - OFF + a), 0 on a clamped axis (cell + a).
template <int OFF>
__device__ __forceinline__ int axis_off(int code) { return code == 2 ? 0 : OFF; }

// The anchor and weights of rel on one axis as the tiles see them: a tile's footprint
// position of stencil node a is anchor - (tile start) + a, for the uniform stencils
// (anchor_axis, nodes anchor - OFF + a) and on a clamped axis (clamped_cell, nodes cell + a).
template <int KIND, int P>
__device__ __forceinline__ int anchor_axis(double rel, int num, int code, double *w);

template <int KIND, int P>
__device__ __forceinline__ int tile_anchor(double rel, int num, int code, double *w) {
  if (KIND == SPLINE && code == 2) {
    int64_t node[4];
    clamped_cell(rel, num, P, 0, node, w);
    return (int)node[0];
  }
  return anchor_axis<KIND, P>(rel, num, code, w);
}

// The anchor (unwrapped) and the weights of rel on one axis; nodes are anchor - OFF + a,
// a = 0..P, OFF = P / 2 (linear: 0), clamped to [0, num - 1] on a bounded axis and taken
// modulo num on a periodic one.  Same arithmetic as mpf_device.h's stencils.
template <int KIND, int P>
__device__ __forceinline__ int anchor_axis(double rel, int num, int code, double *w) {
  const double nd = (double)num;
  double r = rel;
  if (KIND == LINEAR) {
    int l;
    if (code == 1) {
      if (!(r >= 0.0 && r < nd)) r = r - nd * floor(r / nd);
      const int i = (int)fmin(fmax(r, 0.0), nd);  // in range for any input (NaN gives 0)
      l = i >= num ? num - 1 : i;
    } else {
      const int i = (int)fmin(fmax(r, 0.0), nd);
      l = i > num - 2 ? num - 2 : i;
    }
    const double d = r - (double)l;
    const double f = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
    w[0] = 1.0 - f;
    w[1] = f;
    return l;
  }
  if (code == 1) {
    if (!(r >= 0.0 && r < nd)) r = r - nd * floor(r / nd);
  } else {
    if (r < 0.0) r = 0.0;
    if (r > nd - 1.0) r = nd - 1.0;
  }
  const double rc = fmin(fmax(r, 0.0), nd);  // r itself for any finite input
  const int middle = (P % 2 == 0) ? (int)(rc + 0.5) : (int)rc;
  const double o = r - (double)middle;
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
  return middle;
}

// The bounds of rel on one axis within which a marker's stencil needs neither a periodic
// wrap nor a wall clamp: [0, num) on a periodic axis, [0, num - 1] on a bounded one; none
// on a clamped axis (hi < 0).
struct AxisRange {
  double hi[3];
};

__device__ __forceinline__ AxisRange axis_ranges(const Grid &g) {
  AxisRange r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double nd = (double)g.num[d];
    r.hi[d] = g.code[d] == 1 ? nd : (g.code[d] == 0 ? nd - 1.0 : -1.0);
  }
  return r;
}

// Whether anchor_axis's offset of rel from its middle node is rel - cell exactly, for a
// marker binned into stencil cell `cell` (in [0, num)): the position needs neither a
// periodic wrap nor a wall clamp (0 <= rel <= hi; rel < num follows on a periodic axis once
// the offset is in range), and the middle node the binning found is not the wrapped node
// num.  Linear bounded axes always qualify: the left node is the bin's cell for any rel and
// the fraction is clamped.
template <int KIND, int P>
__device__ __forceinline__ bool offset_in_cell(double rel, double o, double hi, int code) {
  if (KIND == LINEAR) return (code == 0) | ((rel >= 0.0) & (rel < hi) & (o >= 0.0) & (o < 1.0));
  const bool near = (P % 2 == 0) ? (fabs(o) <= 0.5) : ((o >= 0.0) & (o < 1.0));
  return (rel >= 0.0) & (rel <= hi) & near;
}

// anchor_axis's weights for a marker binned into stencil cell `cell`: from the offset
// rel - cell where offset_in_cell holds (the same float64 operations, so the same bits),
// from the full selection elsewhere.
template <int KIND, int P>
__device__ __forceinline__ void offset_weights_f64(double o, double *w);

template <int KIND, int P>
__device__ __forceinline__ void cell_weights_f64(double rel, double cell, double hi, int num, int code, double *w) {
  if (KIND == SPLINE && code == 2) {  // clamped knots: the direct kernels' Cox-de Boor weights
    int64_t node[4];
    clamped_cell(rel, num, P, 0, node, w);
    return;
  }
  const double o = rel - cell;
  if (!offset_in_cell<KIND, P>(rel, o, hi, code)) {
    anchor_axis<KIND, P>(rel, num, code, w);
    return;
  }
  offset_weights_f64<KIND, P>(o, w);
}

// Whether a marker binned into stencil cell `cell` takes offset_weights_f64 on every axis
// (no clamped axis, offset_in_cell on each): the test of cell_weights_f64 for all three
// axes at once, so that a caller computes the offset weights unconditionally and takes the
// full selection under one branch that the markers of a binned batch almost never enter.
template <int KIND, int P>
__device__ __forceinline__ bool offsets_in_cell(const double (&rel)[3], const double (&o)[3], const AxisRange &ar,
                                                const Grid &g) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    ok &= !(KIND == SPLINE && g.code[d] == 2) & offset_in_cell<KIND, P>(rel[d], o[d], ar.hi[d], g.code[d]);
  return ok;
}

// The weights of cell_weights_f64 where offset_in_cell holds: a polynomial in the offset.
template <int KIND, int P>
__device__ __forceinline__ void offset_weights_f64(double o, double *w) {
  if (KIND == LINEAR) {
    const double f = o < 0.0 ? 0.0 : (o > 1.0 ? 1.0 : o);
    w[0] = 1.0 - f;
    w[1] = f;
  } else if (P == 0) {
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

template <int KIND, int P>
struct StencilShape {
  static constexpr int S = KIND == LINEAR ? 2 : P + 1;  // nodes per axis
  static constexpr int OFF = KIND == LINEAR ? 0 : P / 2;
};

__device__ __forceinline__ int64_t wrap_index(int64_t k, int64_t num, int code) {
  if (code == 1) return k >= num ? k - num : (k < 0 ? k + num : k);
  return k < 0 ? 0 : (k > num - 1 ? num - 1 : k);
}

// ---- binning ---------------------------------------------------------------------------
//
// 3D markers are binned by anchor cell in tile order: the cells of a tile of
// TILE_X x TILE_Y x TILE_Z (x-major, z fastest) are consecutive bins and tiles follow each
// other in x-major order, so the markers of a tile (and of each x-slab of it) are one
// contiguous run of the binned arrays.  2D markers likewise in tiles of TILE2 x TILE2
// cells (x-major, y fastest).  Partial tiles at the grid's end keep their empty bins.  1D
// markers are binned in cell order.

// Bins of a grid (before the dead bin): tile-padded cells in 2D and 3D, cells in 1D.
__host__ __device__ __forceinline__ int64_t bins_of(int ndim, const int64_t *num) {
  if (ndim == 3) {
    const int64_t ntx = (num[0] + TILE_X - 1) / TILE_X, nty = (num[1] + TILE_Y - 1) / TILE_Y,
                  ntz = (num[2] + TILE_Z - 1) / TILE_Z;
    return ntx * nty * ntz * TILE_CELLS;
  }
  if (ndim == 2) return ((num[0] + TILE2 - 1) / TILE2) * ((num[1] + TILE2 - 1) / TILE2) * TILE2_CELLS;
  return num[0];
}

template <class T, int KIND, int P, int NDIM>
__device__ __forceinline__ uint32_t marker_key(const T *__restrict__ pos, int64_t ps, const T *__restrict__ val,
                                               int64_t vs, int ncomp, int64_t m, const Grid &g, uint32_t dead_key,
                                               double &vmax) {
  // the positions are loaded before the values decide whether they are needed, so that the
  // loads go out together
  T x[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) x[d] = pos[d * ps + m];
  bool live = ncomp == 0;
  for (int c = 0; c < ncomp; ++c) {
    const double v = fabs((double)val[c * vs + m]);
    live |= v != 0.0;
    if (v > vmax && v < 1.0e300) vmax = v;
  }
  if (!live) return dead_key;
  int cell[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    double w[4];
    const double rel = ((double)x[d] - g.lower[d]) * g.inverse[d];
    const int num = (int)g.num[d];
    if (KIND == SPLINE && g.code[d] == 2) {
      cell[d] = clamped_axis_cell(rel, num);  // in [0, num - 2]
      continue;
    }
    int a = anchor_axis<KIND, P>(rel, num, g.code[d], w);
    if (g.code[d] == 1 && a >= num) a -= num;
    cell[d] = a < 0 ? 0 : (a > num - 1 ? num - 1 : a);  // a key in range whatever the position
  }
  if constexpr (NDIM == 3) {
    const int64_t nty = (g.num[1] + TILE_Y - 1) / TILE_Y, ntz = (g.num[2] + TILE_Z - 1) / TILE_Z;
    const int64_t tile = ((int64_t)(cell[0] / TILE_X) * nty + cell[1] / TILE_Y) * ntz + cell[2] / TILE_Z;
    return (uint32_t)(tile * TILE_CELLS + ((cell[0] % TILE_X) * TILE_Y + cell[1] % TILE_Y) * TILE_Z + cell[2] % TILE_Z);
  } else if constexpr (NDIM == 2) {
    const int64_t nty = (g.num[1] + TILE2 - 1) / TILE2;
    const int64_t tile = (int64_t)(cell[0] / TILE2) * nty + cell[1] / TILE2;
    return (uint32_t)(tile * TILE2_CELLS + (cell[0] % TILE2) * TILE2 + cell[1] % TILE2);
  } else {
    return (uint32_t)cell[0];
  }
}

// The counting sort runs in three passes without a global atomic per marker.  Keys are
// split into groups of 2^shift consecutive keys (shift chosen so that a group holds a few
// thousand markers on average).  Pass 1: each of B workgroups takes one contiguous chunk
// of markers, writes their keys and counts its markers per group in LDS; the counts land
// in hist[group][block].  An exclusive scan of hist (group-major) gives every (group,
// block) pair its slot range.  Pass 2: each workgroup walks its chunk again, ranks its
// markers within their group with LDS atomics and writes each marker as one packed record
// (key, index, positions, values) at its slot, so that a marker costs one scattered line
// write rather than one per array.  Pass 3: one workgroup per group holds the group's
// records in registers, counts them per key in LDS, scans, writes the bin offsets of its
// keys, ranks every marker within its bin and writes the sorted arrays one row at a time
// through an LDS stage, so that the global writes are contiguous.  The order of markers
// within one bin is whatever the LDS atomics give; the deposit and the gather do not
// depend on it.
constexpr int BIN_THREADS = 1024;      // workgroup of passes 1, 2 and 3
constexpr int BIN_GROUPS_CAP = 24576;  // LDS counters of passes 1 and 2 (96 KiB)
constexpr int BIN_MAX_BLOCKS = 256;    // workgroups of passes 1 and 2: about one per CU
constexpr int BIN_UNROLL = 4;          // markers per lane per loop trip (loads issued together)
constexpr int BIN_PER_LANE = 8;        // markers a pass-3 lane holds in registers
constexpr int BIN_STAGE = BIN_THREADS * BIN_PER_LANE;  // the largest group pass 3 stages
constexpr int BIN_KEYS_CAP = 4096;     // keys per group of the staged pass 3 (shift <= 12)

// 32-bit words of one packed record: key, index, ndim positions, ncomp values (even, so
// that float64 fields are 8-byte aligned).
__host__ __device__ __forceinline__ int record_words(int ndim, int ncomp, int elem) {
  const int w = 2 + (ndim + ncomp) * (elem / 4);
  return (w + 1) & ~1;
}

// A record of known shape held in registers as its 32-bit words, stored and loaded with
// 16-byte (or, for a length not a multiple of four words, 8-byte) accesses: two stores of a
// float32 3D record with 3 values instead of seven scattered ones.  The records start at a
// 256-byte aligned buffer, so record i is aligned to its own size's power of two.
template <class T, int NDIM, int NCOMP>
struct RecordShape {
  static constexpr int F = NDIM + NCOMP;
  static constexpr int W = (2 + F * (int)(sizeof(T) / 4) + 1) & ~1;  // record_words
};

template <int W>
__device__ __forceinline__ void store_record(uint32_t *dst, const uint32_t (&w)[W]) {
  if constexpr (W % 4 == 0) {
#pragma unroll
    for (int i = 0; i < W / 4; ++i)
      reinterpret_cast<uint4 *>(dst)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < W / 2; ++i) reinterpret_cast<uint2 *>(dst)[i] = make_uint2(w[2 * i], w[2 * i + 1]);
  }
}

template <int W>
__device__ __forceinline__ void load_record(const uint32_t *src, uint32_t (&w)[W]) {
  if constexpr (W % 4 == 0) {
#pragma unroll
    for (int i = 0; i < W / 4; ++i) {
      const uint4 q = reinterpret_cast<const uint4 *>(src)[i];
      w[4 * i] = q.x;
      w[4 * i + 1] = q.y;
      w[4 * i + 2] = q.z;
      w[4 * i + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < W / 2; ++i) {
      const uint2 q = reinterpret_cast<const uint2 *>(src)[i];
      w[2 * i] = q.x;
      w[2 * i + 1] = q.y;
    }
  }
}

template <class T>
__device__ __forceinline__ void put_value(uint32_t *w, T v) {
  if constexpr (sizeof(T) == 4) {
    w[0] = __float_as_uint((float)v);
  } else {
    const unsigned long long b = (unsigned long long)__double_as_longlong((double)v);
    w[0] = (uint32_t)b;
    w[1] = (uint32_t)(b >> 32);
  }
}

template <class T>
__device__ __forceinline__ T get_value(const uint32_t *w) {
  if constexpr (sizeof(T) == 4)
    return (T)__uint_as_float(w[0]);
  else
    return (T)__longlong_as_double((long long)((unsigned long long)w[0] | ((unsigned long long)w[1] << 32)));
}

template <class T, int KIND, int P, int NDIM>
__global__ void __launch_bounds__(BIN_THREADS) bin_histogram_kernel(
    const T *__restrict__ pos, int64_t ps, const T *__restrict__ val, int64_t vs, int64_t n, int ncomp, Grid g,
    uint32_t dead_key, int shift, int groups, int64_t chunk, uint32_t *__restrict__ key, uint32_t *__restrict__ hist,
    unsigned long long *__restrict__ maxbits) {
  __shared__ uint32_t count[BIN_GROUPS_CAP];
  const int t = threadIdx.x, B = gridDim.x, b = blockIdx.x;
  for (int i = t; i < groups; i += BIN_THREADS) count[i] = 0;
  __syncthreads();
  const int64_t begin = (int64_t)b * chunk, end = begin + chunk < n ? begin + chunk : n;
  double vmax = 0.0;
  for (int64_t base = begin; base < end; base += (int64_t)BIN_THREADS * BIN_UNROLL) {
    uint32_t k[BIN_UNROLL];
#pragma unroll
    for (int u = 0; u < BIN_UNROLL; ++u) {
      const int64_t m = base + u * BIN_THREADS + t;
      k[u] = m < end ? marker_key<T, KIND, P, NDIM>(pos, ps, val, vs, ncomp, m, g, dead_key, vmax) : 0u;
    }
#pragma unroll
    for (int u = 0; u < BIN_UNROLL; ++u) {
      const int64_t m = base + u * BIN_THREADS + t;
      if (m < end) {
        key[m] = k[u];
        atomicAdd(&count[k[u] >> shift], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = t; i < groups; i += BIN_THREADS) hist[(int64_t)i * B + b] = count[i];
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(vmax, off, 64);
    vmax = o > vmax ? o : vmax;
  }
  if ((t & 63) == 0 && vmax > 0.0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(vmax));
}

// NDIM / NCOMP fixed: every load of BIN_UNROLL markers is issued before their LDS ranks
// are taken; NCOMP < 0: any ndim / ncomp, loads after the rank.
template <class T, int NDIM, int NCOMP>
__global__ void __launch_bounds__(BIN_THREADS) bin_group_kernel(
    const uint32_t *__restrict__ key, const T *__restrict__ pos, int64_t ps, int ndim, const T *__restrict__ val,
    int64_t vs, int ncomp, int64_t n, int shift, int groups, int64_t chunk, const uint32_t *__restrict__ slots,
    uint32_t *__restrict__ records) {
  constexpr int F = NCOMP >= 0 ? NDIM + NCOMP : 1;
  __shared__ uint32_t next[BIN_GROUPS_CAP];
  const int t = threadIdx.x, B = gridDim.x, b = blockIdx.x;
  const int stride = record_words(ndim, ncomp, (int)sizeof(T));
  for (int i = t; i < groups; i += BIN_THREADS) next[i] = slots[(int64_t)i * B + b];
  __syncthreads();
  const int64_t begin = (int64_t)b * chunk, end = begin + chunk < n ? begin + chunk : n;
  for (int64_t base = begin; base < end; base += (int64_t)BIN_THREADS * BIN_UNROLL) {
    uint32_t k[BIN_UNROLL];
    T f[BIN_UNROLL][F];
#pragma unroll
    for (int u = 0; u < BIN_UNROLL; ++u) {
      const int64_t m = base + u * BIN_THREADS + t;
      const bool in = m < end;
      k[u] = in ? key[m] : 0u;
      if constexpr (NCOMP >= 0) {
#pragma unroll
        for (int d = 0; d < NDIM; ++d) f[u][d] = in ? pos[d * ps + m] : T(0);
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) f[u][NDIM + c] = in ? val[c * vs + m] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < BIN_UNROLL; ++u) {
      const int64_t m = base + u * BIN_THREADS + t;
      if (m < end) {
        const uint32_t dst = atomicAdd(&next[k[u] >> shift], 1u);
        uint32_t *r = records + (int64_t)dst * stride;
        if constexpr (NCOMP >= 0) {
          using R = RecordShape<T, NDIM, NCOMP>;
          uint32_t w[R::W];
          w[0] = k[u];
          w[1] = (uint32_t)m;
#pragma unroll
          for (int q = 0; q < R::F; ++q) put_value<T>(w + 2 + q * (int)(sizeof(T) / 4), f[u][q]);
          if constexpr (2 + R::F * (int)(sizeof(T) / 4) < R::W) w[R::W - 1] = 0u;
          store_record<R::W>(r, w);
        } else {
          *(unsigned long long *)r = (unsigned long long)k[u] | ((unsigned long long)m << 32);
          T *out = (T *)(r + 2);
          for (int d = 0; d < ndim; ++d) out[d] = pos[d * ps + m];
          for (int c = 0; c < ncomp; ++c) out[ndim + c] = val[c * vs + m];
        }
      }
    }
  }
}

__device__ __forceinline__ unsigned long long wave_inclusive_sum(unsigned long long v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// In-place exclusive scan of a[0..count) by a workgroup of BLOCK lanes (each lane scans a
// run of consecutive entries).  Ends with a barrier.
template <int BLOCK>
__device__ __forceinline__ void block_exclusive_scan(uint32_t *a, int count, uint32_t *wave_sums) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int per = (count + BLOCK - 1) / BLOCK, lo = t * per, hi = lo + per < count ? lo + per : count;
  uint32_t sum = 0;
  for (int i = lo; i < hi; ++i) sum += a[i];
  unsigned long long incl = wave_inclusive_sum(sum);
  if (lane == 63) wave_sums[w] = (uint32_t)incl;
  __syncthreads();
  uint32_t before = (uint32_t)incl - sum;
  for (int v = 0; v < w; ++v) before += wave_sums[v];
  for (int i = lo; i < hi; ++i) {
    const uint32_t c = a[i];
    a[i] = before;
    before += c;
  }
  __syncthreads();
}

// Pass 3 for one group whose markers are records [lo, hi): counts per key into cnt
// (keys entries, zero on entry), scans and writes start[] for the group's keys.  On return
// cnt holds each key's first slot relative to lo.
__device__ __forceinline__ void group_offsets(const uint32_t *__restrict__ records, int stride, uint32_t lo,
                                              uint32_t hi, uint32_t first_key, int keys, uint32_t *cnt,
                                              uint32_t *wave_sums, uint32_t *__restrict__ start) {
  for (uint32_t i = lo + threadIdx.x; i < hi; i += BIN_THREADS)
    atomicAdd(&cnt[records[(int64_t)i * stride] - first_key], 1u);
  __syncthreads();
  block_exclusive_scan<BIN_THREADS>(cnt, keys, wave_sums);
  for (int i = threadIdx.x; i < keys; i += BIN_THREADS) start[first_key + i] = lo + cnt[i];
  __syncthreads();  // every offset read for start[] before the placement moves them on
}

// Placement straight from the records (any group size, any ndim / ncomp).
template <class T>
__device__ __forceinline__ void place_direct(const uint32_t *__restrict__ records, int stride, uint32_t lo,
                                             uint32_t hi, uint32_t first_key, uint32_t *cnt, int ndim, int ncomp,
                                             int64_t n, T *__restrict__ spos, T *__restrict__ sval,
                                             int64_t *__restrict__ order) {
  for (uint32_t i = lo + threadIdx.x; i < hi; i += BIN_THREADS) {
    const uint32_t *r = records + (int64_t)i * stride;
    const uint32_t dst = lo + atomicAdd(&cnt[r[0] - first_key], 1u);
    order[dst] = (int64_t)r[1];
    const T *f = (const T *)(r + 2);
    for (int d = 0; d < ndim; ++d) spos[d * n + dst] = f[d];
    for (int c = 0; c < ncomp; ++c) sval[c * n + dst] = f[ndim + c];
  }
}

// The group range and key range of pass-3 workgroup grp.
__device__ __forceinline__ void group_range(const uint32_t *__restrict__ slots, int grp, int groups, int blocks,
                                            int64_t n, int shift, uint32_t dead_key, uint32_t &lo, uint32_t &hi,
                                            uint32_t &first_key, int &keys) {
  lo = slots[(int64_t)grp * blocks];
  hi = grp + 1 < groups ? slots[(int64_t)(grp + 1) * blocks] : (uint32_t)n;
  first_key = (uint32_t)grp << shift;
  const uint32_t left = dead_key + 1 - first_key;
  keys = (int)(left < (1u << shift) ? left : (1u << shift));
}

// Staged pass 3: the group's records in registers (BIN_PER_LANE per lane), each output
// row permuted in LDS and written contiguously.  Groups larger than BIN_STAGE take the
// direct placement.
template <class T, int NDIM, int NCOMP>
__global__ void __launch_bounds__(BIN_THREADS) bin_cells_staged(
    const uint32_t *__restrict__ records, int64_t n, int shift, int groups, int blocks, uint32_t dead_key,
    const uint32_t *__restrict__ slots, T *__restrict__ spos, T *__restrict__ sval, int64_t *__restrict__ order,
    uint32_t *__restrict__ start) {
  constexpr int F = NDIM + NCOMP;
  __shared__ uint32_t cnt[BIN_KEYS_CAP];
  __shared__ T stage[BIN_STAGE];
  __shared__ uint32_t wave_sums[BIN_THREADS / 64];
  const int t = threadIdx.x, stride = record_words(NDIM, NCOMP, (int)sizeof(T));
  uint32_t lo, hi, first_key;
  int keys;
  group_range(slots, blockIdx.x, groups, blocks, n, shift, dead_key, lo, hi, first_key, keys);
  for (int i = t; i < keys; i += BIN_THREADS) cnt[i] = 0;
  if (blockIdx.x == groups - 1 && t == 0) start[dead_key + 1] = (uint32_t)n;
  __syncthreads();
  if (hi - lo > (uint32_t)BIN_STAGE) {
    group_offsets(records, stride, lo, hi, first_key, keys, cnt, wave_sums, start);
    place_direct<T>(records, stride, lo, hi, first_key, cnt, NDIM, NCOMP, n, spos, sval, order);
    return;
  }
  uint32_t k[BIN_PER_LANE], idx[BIN_PER_LANE];
  T f[BIN_PER_LANE][F > 0 ? F : 1];
#pragma unroll
  for (int j = 0; j < BIN_PER_LANE; ++j) {
    const uint32_t i = lo + t + j * BIN_THREADS;
    k[j] = 0xFFFFFFFFu;
    if (i < hi) {
      using R = RecordShape<T, NDIM, NCOMP>;
      uint32_t w[R::W];
      load_record<R::W>(records + (int64_t)i * stride, w);
      k[j] = w[0] - first_key;
      idx[j] = w[1];
#pragma unroll
      for (int q = 0; q < F; ++q) f[j][q] = get_value<T>(w + 2 + q * (int)(sizeof(T) / 4));
    }
  }
#pragma unroll
  for (int j = 0; j < BIN_PER_LANE; ++j)
    if (k[j] != 0xFFFFFFFFu) atomicAdd(&cnt[k[j]], 1u);
  __syncthreads();
  block_exclusive_scan<BIN_THREADS>(cnt, keys, wave_sums);
  for (int i = t; i < keys; i += BIN_THREADS) start[first_key + i] = lo + cnt[i];
  __syncthreads();
  uint32_t dst[BIN_PER_LANE];
#pragma unroll
  for (int j = 0; j < BIN_PER_LANE; ++j) dst[j] = k[j] != 0xFFFFFFFFu ? atomicAdd(&cnt[k[j]], 1u) : 0u;
  const int size = (int)(hi - lo);
  // the input index, as the permutation
  uint32_t *words = (uint32_t *)stage;
#pragma unroll
  for (int j = 0; j < BIN_PER_LANE; ++j)
    if (k[j] != 0xFFFFFFFFu) words[dst[j]] = idx[j];
  __syncthreads();
  for (int i = t; i < size; i += BIN_THREADS) order[lo + i] = (int64_t)words[i];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < F; ++q) {
#pragma unroll
    for (int j = 0; j < BIN_PER_LANE; ++j)
      if (k[j] != 0xFFFFFFFFu) stage[dst[j]] = f[j][q];
    __syncthreads();
    T *row = q < NDIM ? spos + q * n : sval + (q - NDIM) * n;
    for (int i = t; i < size; i += BIN_THREADS) row[lo + i] = stage[i];
    __syncthreads();
  }
}

// Pass 3 for any ndim / ncomp and any group size (direct placement).
template <class T, int KEYS>
__global__ void __launch_bounds__(BIN_THREADS) bin_cells_direct(
    const uint32_t *__restrict__ records, int ndim, int ncomp, int64_t n, int shift, int groups, int blocks,
    uint32_t dead_key, const uint32_t *__restrict__ slots, T *__restrict__ spos, T *__restrict__ sval,
    int64_t *__restrict__ order, uint32_t *__restrict__ start) {
  __shared__ uint32_t cnt[KEYS];
  __shared__ uint32_t wave_sums[BIN_THREADS / 64];
  const int t = threadIdx.x, stride = record_words(ndim, ncomp, (int)sizeof(T));
  uint32_t lo, hi, first_key;
  int keys;
  group_range(slots, blockIdx.x, groups, blocks, n, shift, dead_key, lo, hi, first_key, keys);
  for (int i = t; i < keys; i += BIN_THREADS) cnt[i] = 0;
  if (blockIdx.x == groups - 1 && t == 0) start[dead_key + 1] = (uint32_t)n;
  __syncthreads();
  group_offsets(records, stride, lo, hi, first_key, keys, cnt, wave_sums, start);
  place_direct<T>(records, stride, lo, hi, first_key, cnt, ndim, ncomp, n, spos, sval, order);
}

// ---- re-binning markers kept in bins order ----------------------------------------------
//
// A particle code that keeps its markers in the bins' order moves them and bins them again:
// most stay in their 3D tile and most of the others reach a neighbouring tile.  Pass 1 (one
// workgroup per old tile, and one for the old dead bin) computes every marker's new key
// (marker_key, so keys and the value maximum are those a fresh binning finds), counts the
// markers of each destination tile (a global atomic per destination tile and workgroup) and
// lists the markers that leave for a neighbouring tile, grouped by destination, in the old
// tile's own range of a mover array.  Markers whose values became all zero go to one dead list; markers that go further
// are counted as far, and the call then sorts from scratch.  An exclusive scan of the tile
// counts gives every new tile its first slot.  Pass 2 (one workgroup per new tile, and one
// for the dead bin) visits the tile's markers -- those that stayed, from its old range, and
// those that came, from the parts of its neighbours' mover lists sent to it -- counts them
// per cell in LDS, writes the cells' offsets, ranks each marker within its cell and copies
// the rows in the new order through that permutation.  The order within a cell is what
// the LDS atomics give, as in the full sort.
constexpr int REBIN_THREADS = 256;       // pass 1 and the dead bin
constexpr int REBIN_PLACE_THREADS = 512;  // pass 2

struct TileGrid {
  int nt[3];    // tiles per axis
  int code[3];  // 1 periodic
  int tiles;    // nt[0] nt[1] nt[2]; the dead bin is tile `tiles`
};

__device__ __forceinline__ void tile_coords(const TileGrid &tg, int tile, int (&c)[3]) {
  c[2] = tile % tg.nt[2];
  c[1] = (tile / tg.nt[2]) % tg.nt[1];
  c[0] = tile / (tg.nt[2] * tg.nt[1]);
}

// The offsets along axis d from tile coordinate c that reach distinct tiles, 0 first:
// a periodic axis of two tiles has +1 only (it is also -1), one of one tile none; a bounded
// axis stops at its ends.
__device__ __forceinline__ int axis_offsets(const TileGrid &tg, int d, int c, int (&off)[3]) {
  int k = 0;
  off[k++] = 0;
  if (tg.code[d] == 1) {
    if (tg.nt[d] >= 2) off[k++] = 1;
    if (tg.nt[d] >= 3) off[k++] = -1;
  } else {
    if (c + 1 < tg.nt[d]) off[k++] = 1;
    if (c >= 1) off[k++] = -1;
  }
  return k;
}

__device__ __forceinline__ int axis_step(const TileGrid &tg, int d, int c, int o) {
  int v = c + o;
  if (tg.code[d] == 1) v = v < 0 ? v + tg.nt[d] : (v >= tg.nt[d] ? v - tg.nt[d] : v);
  return v;
}

__device__ __forceinline__ int tile_at(const TileGrid &tg, int x, int y, int z) {
  return (x * tg.nt[1] + y) * tg.nt[2] + z;
}

// Tile u's place 0..26 among tile t's neighbours ((o0 + 1) 9 + (o1 + 1) 3 + o2 + 1 for the
// offsets axis_offsets gives), or -1 when u is further.
__device__ __forceinline__ int neighbour_slot(const TileGrid &tg, int t, int u) {
  int a[3], b[3];
  tile_coords(tg, t, a);
  tile_coords(tg, u, b);
  int slot = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int o = b[d] - a[d];
    if (tg.code[d] == 1) {
      if (o > 1) o -= tg.nt[d];
      if (o < -1) o += tg.nt[d];
      if (tg.nt[d] == 2 && o == -1) o = 1;
    }
    if (o < -1 || o > 1) return -1;
    slot = slot * 3 + o + 1;
  }
  return slot;
}

// The old dead bin is split over REBIN_DEAD_GROUPS workgroups of pass 1, and the new one is
// filled by as many workgroups of pass 3 (its markers need no order).
constexpr int REBIN_DEAD_GROUPS = 64;

// Lane-aggregated LDS counter: every lane with `take` gets a distinct slot of *counter,
// one atomic per wavefront.
__device__ __forceinline__ uint32_t wave_take(uint32_t *counter, bool take) {
  const unsigned long long mask = __ballot(take);
  if (!mask) return 0u;
  const int lane = threadIdx.x & 63, first = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, first, 64);
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Pass 1 also writes each tile's mover list grouped by destination: moff[tile * 28 + s]
// is where the markers for its neighbour s (0..26, neighbour_slot) start within the list,
// moff[tile * 28 + 27] its length.
constexpr int REBIN_MOFF = 28;
constexpr int REBIN_UNROLL = 4;  // markers per work-item and trip, their loads issued together
constexpr int REBIN_MSTAGE = 1024;  // movers of a tile that pass 1 keeps in LDS

template <class T, int KIND, int P>
__global__ void __launch_bounds__(REBIN_THREADS) rebin_count_kernel(
    const T *__restrict__ pos, int64_t ps, const T *__restrict__ val, int64_t vs, int ncomp, Grid g, TileGrid tg,
    const uint32_t *__restrict__ ostart, uint32_t *__restrict__ keys, uint2 *__restrict__ movers,
    uint32_t *__restrict__ moff, uint32_t *__restrict__ total, uint32_t *__restrict__ dead_list,
    uint32_t *__restrict__ counters, unsigned long long *__restrict__ maxbits) {
  __shared__ uint32_t dest[27], doff[28];
  __shared__ uint32_t stayed, died, moved;
  __shared__ uint2 staged[REBIN_MSTAGE];  // the movers (index, key) as found
  __shared__ uint8_t staged_slot[REBIN_MSTAGE];
  const int t = threadIdx.x;
  const uint32_t dead_key = (uint32_t)tg.tiles * TILE_CELLS;
  // workgroups past the tiles take one piece each of the old dead bin
  const int item = (int)blockIdx.x;
  const bool dead_part = item >= tg.tiles;
  const int tile = dead_part ? tg.tiles : item;
  if (t < 27) dest[t] = 0;
  if (t == 0) stayed = died = moved = 0;
  __syncthreads();
  uint32_t lo = ostart[(int64_t)tile * TILE_CELLS];
  uint32_t hi = ostart[dead_part ? (int64_t)dead_key + 1 : (int64_t)(tile + 1) * TILE_CELLS];
  if (dead_part) {
    const uint32_t piece = (hi - lo + REBIN_DEAD_GROUPS - 1) / REBIN_DEAD_GROUPS, k = item - tg.tiles;
    const uint32_t a = lo + k * piece;
    lo = a < hi ? a : hi;
    hi = lo + piece < hi ? lo + piece : hi;
  }
  double vmax = 0.0;
  // whole wavefronts go round the loop together (the lane-aggregated counters need them)
  for (uint32_t base = lo; base < hi; base += REBIN_THREADS * REBIN_UNROLL) {
    uint32_t key[REBIN_UNROLL];
#pragma unroll
    for (int q = 0; q < REBIN_UNROLL; ++q) {
      const uint32_t m = base + q * REBIN_THREADS + t;
      key[q] = m < hi ? marker_key<T, KIND, P, 3>(pos, ps, val, vs, ncomp, m, g, dead_key, vmax) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int q = 0; q < REBIN_UNROLL; ++q) {
      const uint32_t m = base + q * REBIN_THREADS + t;
      const bool in = key[q] != 0xFFFFFFFFu;
      if (in) keys[m] = key[q];
      const int u = in ? (int)(key[q] / TILE_CELLS) : -1;
      const bool stay = in && u == tile;  // also a dead marker that stays dead
      const bool die = in && !stay && u == tg.tiles;
      const int s = in && !stay && !die && !dead_part ? neighbour_slot(tg, tile, u) : -1;
      if (in && !stay && !die && s < 0) atomicAdd(&counters[1], 1u);  // further than a neighbouring tile
      if (die) dead_list[atomicAdd(&counters[0], 1u)] = m;
      (void)wave_take(&stayed, stay);
      (void)wave_take(&died, die);
      const uint32_t r = wave_take(&moved, s >= 0);
      if (s >= 0) {
        atomicAdd(&dest[s], 1u);
        if (r < (uint32_t)REBIN_MSTAGE) {
          staged[r] = make_uint2(m, key[q]);
          staged_slot[r] = (uint8_t)s;
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(vmax, off, 64);
    vmax = o > vmax ? o : vmax;
  }
  if ((t & 63) == 0 && vmax > 0.0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(vmax));
  __syncthreads();
  if (t == 0) {
    if (stayed) atomicAdd(&total[tile], stayed);
    if (died) atomicAdd(&total[tg.tiles], died);
    if (!dead_part) {
      doff[0] = 0;
      for (int i = 0; i < 27; ++i) doff[i + 1] = doff[i] + dest[i];
    }
  }
  __syncthreads();
  if (dead_part) return;
  if (t < REBIN_MOFF) moff[(int64_t)tile * REBIN_MOFF + t] = doff[t];
  if (t < 27 && dest[t]) {
    int a[3];
    tile_coords(tg, tile, a);
    const int u = tile_at(tg, axis_step(tg, 0, a[0], t / 9 - 1), axis_step(tg, 1, a[1], (t / 3) % 3 - 1),
                          axis_step(tg, 2, a[2], t % 3 - 1));
    atomicAdd(&total[u], dest[t]);
  }
  if (doff[27] == 0) return;
  // the movers, grouped by destination: from LDS, or (more than REBIN_MSTAGE of them) from
  // the keys again (just written, in cache)
  if (moved <= (uint32_t)REBIN_MSTAGE) {
    for (uint32_t r = t; r < moved; r += REBIN_THREADS)
      movers[lo + atomicAdd(&doff[staged_slot[r]], 1u)] = staged[r];
    return;
  }
  for (uint32_t m = lo + t; m < hi; m += REBIN_THREADS) {
    const uint32_t key = keys[m];
    const int u = (int)(key / TILE_CELLS);
    if (u == tile || u == tg.tiles) continue;
    const int s = neighbour_slot(tg, tile, u);
    if (s >= 0) movers[lo + atomicAdd(&doff[s], 1u)] = make_uint2(m, key);
  }
}

// The sources of new tile u's markers as one index range: its old range (the markers that
// stayed, among others), then the lists of markers its distinct neighbours sent to it, in LDS.
struct RebinSources {
  uint32_t lo, own;             // the old range of tile u
  int count;                    // neighbours
  uint32_t first[26], off[27];  // list start; prefix of the list lengths
};

__device__ __forceinline__ void rebin_sources(const TileGrid &tg, int u, const uint32_t *__restrict__ ostart,
                                              const uint32_t *__restrict__ moff, RebinSources &src) {
  const int t = threadIdx.x;
  int a[3], ox[3], oy[3], oz[3];
  tile_coords(tg, u, a);
  const int nx = axis_offsets(tg, 0, a[0], ox), ny = axis_offsets(tg, 1, a[1], oy), nz = axis_offsets(tg, 2, a[2], oz);
  const int k = nx * ny * nz - 1;  // the neighbours (offset 0 0 0 is the first combination)
  if (t == 0) {
    src.lo = ostart[(int64_t)u * TILE_CELLS];
    src.own = ostart[(int64_t)(u + 1) * TILE_CELLS] - src.lo;
    src.count = k;
  }
  if (t < k) {
    const int c = t + 1, i = c / (ny * nz), j = (c / nz) % ny, l = c % nz;
    const int nb = tile_at(tg, axis_step(tg, 0, a[0], ox[i]), axis_step(tg, 1, a[1], oy[j]), axis_step(tg, 2, a[2], oz[l]));
    const int slot = neighbour_slot(tg, nb, u);  // u as neighbour nb sees it
    const uint32_t *mo = moff + (int64_t)nb * REBIN_MOFF + slot;
    src.first[t] = ostart[(int64_t)nb * TILE_CELLS] + mo[0];
    src.off[t + 1] = mo[1] - mo[0];
  }
  __syncthreads();
  if (t == 0) {
    src.off[0] = 0;
    for (int i = 1; i <= k; ++i) src.off[i] += src.off[i - 1];
  }
  __syncthreads();
}

// Calls f(m, key) for every marker of new tile u among its sources, REBIN_UNROLL candidates
// per work-item and trip with their loads issued together.
template <class F>
__device__ __forceinline__ void rebin_visit(int u, const RebinSources &src, const uint32_t *__restrict__ keys,
                                            const uint2 *__restrict__ movers, F &&f) {
  const uint32_t total = src.own + src.off[src.count];
  for (uint32_t base = 0; base < total; base += REBIN_PLACE_THREADS * REBIN_UNROLL) {
    uint32_t m[REBIN_UNROLL], key[REBIN_UNROLL];
#pragma unroll
    for (int q = 0; q < REBIN_UNROLL; ++q) {
      const uint32_t e = base + q * REBIN_PLACE_THREADS + threadIdx.x;
      m[q] = 0xFFFFFFFFu;
      key[q] = 0xFFFFFFFFu;
      if (e < src.own) {
        m[q] = src.lo + e;
        key[q] = keys[m[q]];
      } else if (e < total) {
        const uint32_t j = e - src.own;
        int lo = 0, hi = src.count;  // the list holding entry j: off[lo] <= j < off[lo + 1]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (src.off[mid] <= j) lo = mid;
          else hi = mid;
        }
        const uint2 mv = movers[src.first[lo] + (j - src.off[lo])];
        m[q] = mv.x;
        key[q] = mv.y;
      }
    }
#pragma unroll
    for (int q = 0; q < REBIN_UNROLL; ++q)
      if (key[q] != 0xFFFFFFFFu && (int)(key[q] / TILE_CELLS) == u) f(m[q], key[q]);
  }
}

template <class T>
__device__ __forceinline__ void rebin_copy(const T *__restrict__ pos, int64_t ps, const T *__restrict__ val,
                                           int64_t vs, int ncomp, int64_t n, const int64_t *__restrict__ oorder,
                                           uint32_t m, uint32_t slot, T *__restrict__ spos, T *__restrict__ sval,
                                           int64_t *__restrict__ order) {
#pragma unroll
  for (int d = 0; d < 3; ++d) spos[d * n + slot] = pos[d * ps + m];
  for (int c = 0; c < ncomp; ++c) sval[c * n + slot] = val[c * vs + m];
  order[slot] = oorder[m];
}

// Markers of a new tile that pass 2 stages through LDS (larger tiles are placed directly).
constexpr int REBIN_STAGE = 8192;

// NC >= 0: the component count fixed, every load of two markers issued before their stores;
// NC < 0: ncomp at run time.
template <class T, int NC>
__global__ void __launch_bounds__(REBIN_PLACE_THREADS) rebin_place_kernel(
    const T *__restrict__ pos, int64_t ps, const T *__restrict__ val, int64_t vs, int ncomp, int64_t n, TileGrid tg,
    const uint32_t *__restrict__ ostart, const int64_t *__restrict__ oorder, const uint32_t *__restrict__ keys,
    const uint2 *__restrict__ movers, const uint32_t *__restrict__ moff, const uint32_t *__restrict__ tstart,
    T *__restrict__ spos, T *__restrict__ sval, int64_t *__restrict__ order, uint32_t *__restrict__ start) {
  __shared__ uint32_t cnt[TILE_CELLS];
  __shared__ uint32_t perm[REBIN_STAGE];  // the old index of the tile's i-th new marker
  __shared__ uint32_t wave_sums[REBIN_PLACE_THREADS / 64];
  __shared__ RebinSources src;
  const int t = threadIdx.x, u = blockIdx.x;
  for (int i = t; i < TILE_CELLS; i += REBIN_PLACE_THREADS) cnt[i] = 0;
  rebin_sources(tg, u, ostart, moff, src);  // ends with a barrier
  rebin_visit(u, src, keys, movers, [&](uint32_t, uint32_t key) { atomicAdd(&cnt[key % TILE_CELLS], 1u); });
  __syncthreads();
  block_exclusive_scan<REBIN_PLACE_THREADS>(cnt, TILE_CELLS, wave_sums);
  const uint32_t base = tstart[u], size = tstart[u + 1] - base;
  for (int i = t; i < TILE_CELLS; i += REBIN_PLACE_THREADS) start[(int64_t)u * TILE_CELLS + i] = base + cnt[i];
  __syncthreads();  // every offset read before the placement moves them on
  if (size > (uint32_t)REBIN_STAGE) {  // a crowded tile: each marker straight to its slot
    rebin_visit(u, src, keys, movers, [&](uint32_t m, uint32_t key) {
      rebin_copy<T>(pos, ps, val, vs, ncomp, n, oorder, m, base + atomicAdd(&cnt[key % TILE_CELLS], 1u), spos, sval,
                    order);
    });
    return;
  }
  // the permutation in LDS, then the markers copied in order: contiguous writes, and reads
  // within the tile's old range (runs of a cell's markers that stayed together)
  rebin_visit(u, src, keys, movers, [&](uint32_t m, uint32_t key) { perm[atomicAdd(&cnt[key % TILE_CELLS], 1u)] = m; });
  __syncthreads();
  if constexpr (NC >= 0) {
    constexpr int K = 2;
    for (uint32_t i0 = 0; i0 < size; i0 += REBIN_PLACE_THREADS * K) {
      uint32_t m[K];
      int64_t o[K];
      T p[K][3], v[K][NC > 0 ? NC : 1];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const uint32_t i = i0 + q * REBIN_PLACE_THREADS + t;
        m[q] = perm[i < size ? i : size - 1];
        o[q] = oorder[m[q]];
#pragma unroll
        for (int d = 0; d < 3; ++d) p[q][d] = pos[d * ps + m[q]];
#pragma unroll
        for (int c = 0; c < NC; ++c) v[q][c] = val[c * vs + m[q]];
      }
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const uint32_t i = i0 + q * REBIN_PLACE_THREADS + t;
        if (i < size) {
          order[base + i] = o[q];
#pragma unroll
          for (int d = 0; d < 3; ++d) spos[d * n + base + i] = p[q][d];
#pragma unroll
          for (int c = 0; c < NC; ++c) sval[c * n + base + i] = v[q][c];
        }
      }
    }
  } else {
    for (uint32_t i = t; i < size; i += REBIN_PLACE_THREADS) rebin_copy<T>(pos, ps, val, vs, ncomp, n, oorder, perm[i], base + i,
                                                                     spos, sval, order);
  }
}

// The new dead bin: the old dead bin's markers that stayed dead and the dead list, in any
// order, placed through one global cursor (counters[2]).
template <class T>
__global__ void __launch_bounds__(REBIN_THREADS) rebin_dead_kernel(
    const T *__restrict__ pos, int64_t ps, const T *__restrict__ val, int64_t vs, int ncomp, int64_t n, TileGrid tg,
    const uint32_t *__restrict__ ostart, const int64_t *__restrict__ oorder, const uint32_t *__restrict__ keys,
    const uint32_t *__restrict__ tstart, const uint32_t *__restrict__ dead_list, uint32_t *__restrict__ counters,
    T *__restrict__ spos, T *__restrict__ sval, int64_t *__restrict__ order, uint32_t *__restrict__ start) {
  __shared__ uint32_t cursor;
  const int64_t dead_key = (int64_t)tg.tiles * TILE_CELLS;
  const uint32_t base = tstart[tg.tiles], lo = ostart[dead_key], hi = ostart[dead_key + 1], ndead = counters[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    start[dead_key] = base;
    start[dead_key + 1] = (uint32_t)n;
  }
  const uint32_t stride = gridDim.x * REBIN_THREADS;
  for (uint32_t round = blockIdx.x * REBIN_THREADS; round < (hi - lo) + ndead; round += stride) {
    const uint32_t i = round + threadIdx.x;
    uint32_t m = 0;
    bool take = false;
    if (i < hi - lo) {
      m = lo + i;
      take = keys[m] == (uint32_t)dead_key;
    } else if (i < (hi - lo) + ndead) {
      m = dead_list[i - (hi - lo)];
      take = true;
    }
    if (threadIdx.x == 0) cursor = 0;
    __syncthreads();
    const uint32_t r = wave_take(&cursor, take);
    __syncthreads();
    if (threadIdx.x == 0 && cursor) cursor = atomicAdd(&counters[2], cursor);
    __syncthreads();
    if (take) rebin_copy<T>(pos, ps, val, vs, ncomp, n, oorder, m, base + cursor + r, spos, sval, order);
    __syncthreads();
  }
}

// order[i] = old_order[order[i]]: a permutation relative to the previous bins, made
// relative to the markers' first input.
__global__ void compose_order_kernel(int64_t *__restrict__ order, const int64_t *__restrict__ old_order, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    order[i] = old_order[order[i]];
}

// ---- tiled deposit --------------------------------------------------------------------

struct Tiling {
  int tx, ty, tz;          // tile extent in anchor cells (a workgroup's: ZT tiles along z)
  int ntx, nty, ntz;       // workgroup tiles per axis
  int ntz_bins;            // 8 x 8 x 16 bin tiles along z (the bins' key layout)
  int cx, cy, cz;          // colours per axis
  int color_x, color_y, color_z;  // the colour of this launch
  int tiles_x, tiles_y, tiles_z;  // tiles of this colour per axis
};

// The k-th tile of colour `color` along an axis of nt tiles.  Two colours alternate; with
// three, tile 0 has colour 2 to itself and tiles 1 .. nt - 1 alternate (colours_for).
__device__ __forceinline__ int tile_of_color(int k, int colors, int color) {
  if (colors == 1) return k;
  if (colors == 3) return color == 2 ? 0 : 1 + 2 * k + color;
  return 2 * k + color;
}

// ---- work items of a tile ---------------------------------------------------------------
//
// The markers of a cell are cut into at most SWEEPS items of about ITEM markers (sizes
// differ by at most one).  A tile's items are listed part by part (a part is BLOCK
// consecutive cells of the tile, one per work-item), within a part sweep by sweep (the
// first item of every cell, then the second, ...) and within a sweep in cell order.  The
// lanes of a wavefront therefore take items of neighbouring cells: their LDS addresses for
// one stencil node are consecutive (no bank conflicts) and their markers lie close together
// in the binned arrays.  An item is a uint16: cell | sweep << 12.
constexpr int ITEM = 3, SWEEPS = 4;

__device__ __forceinline__ uint32_t items_of(uint32_t count) {
  const uint32_t k = (count + ITEM - 1) / ITEM;
  return k < (uint32_t)SWEEPS ? k : (uint32_t)SWEEPS;
}


// Fill items[] for the cells whose marker offsets are cfirst[0..TC]; returns the number of
// items.  One block-wide scan: each work-item packs its cell's "has a j-th item" flags
// (j < SWEEPS) into 16-bit fields of one 64-bit word, so a single prefix sum ranks the
// cell in every sweep at once.  Ends with a barrier.
template <int BLOCK, int TC>
__device__ uint32_t build_items(const uint32_t *cfirst, uint16_t *items, unsigned long long *wave_sums) {
  constexpr int PARTS = TC / BLOCK, WAVES = BLOCK / 64;
  static_assert(TC % BLOCK == 0 && TC <= 4096 && BLOCK % 64 == 0, "item table");
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  unsigned long long mine[PARTS], incl[PARTS];
  uint32_t nitems[PARTS];
#pragma unroll
  for (int p = 0; p < PARTS; ++p) {
    const int c = p * BLOCK + t;
    nitems[p] = items_of(cfirst[c + 1] - cfirst[c]);
    unsigned long long f = 0;
#pragma unroll
    for (int j = 0; j < SWEEPS; ++j) f |= (unsigned long long)(nitems[p] > (uint32_t)j) << (16 * j);
    mine[p] = f;
    incl[p] = wave_inclusive_sum(f);
    if (lane == 63) wave_sums[w * PARTS + p] = incl[p];
  }
  __syncthreads();
  uint32_t base = 0;
#pragma unroll
  for (int p = 0; p < PARTS; ++p) {
    unsigned long long before = 0, total = 0;
#pragma unroll
    for (int v = 0; v < WAVES; ++v) {
      const unsigned long long s = wave_sums[v * PARTS + p];
      before += v < w ? s : 0ull;
      total += s;
    }
    const unsigned long long rank = before + incl[p] - mine[p];
#pragma unroll
    for (int j = 0; j < SWEEPS; ++j) {
      if (nitems[p] > (uint32_t)j) items[base + (uint32_t)((rank >> (16 * j)) & 0xFFFF)] = (uint16_t)((p * BLOCK + t) | (j << 12));
      base += (uint32_t)((total >> (16 * j)) & 0xFFFF);
    }
  }
  __syncthreads();
  return base;
}

// A loaded value held at this point: its load completes here, and no load before this point
// is moved past it (the compiler otherwise sinks a load into the guarded store that uses it,
// so that a tile's start waits on one load after another).
template <class V>
__device__ __forceinline__ void loaded_here(const V &v) {
  asm volatile("" ::"v"(v) : "memory");
}

// The tile's TC + 1 cell offsets into LDS, every load issued before the first store.
template <int BLOCK, int TC>
__device__ __forceinline__ void load_offsets(const uint32_t *__restrict__ start, uint32_t *cfirst) {
  constexpr int E = (TC + 1 + BLOCK - 1) / BLOCK;
  uint32_t v[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    v[u] = start[i <= TC ? i : TC];
  }
#pragma unroll
  for (int u = 0; u < E; ++u) loaded_here(v[u]);
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    if (i <= TC) cfirst[i] = v[u];
  }
}

// The offsets of a workgroup tile's first `valid` cells (valid <= TC); the entries past them
// repeat the last valid one (empty cells), so no offset beyond the bins is read.
template <int BLOCK, int TC>
__device__ __forceinline__ void load_offsets_valid(const uint32_t *__restrict__ start, int valid, uint32_t *cfirst) {
  constexpr int E = (TC + 1 + BLOCK - 1) / BLOCK;
  uint32_t v[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    v[u] = start[i <= valid ? i : valid];
  }
#pragma unroll
  for (int u = 0; u < E; ++u) loaded_here(v[u]);
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    if (i <= TC) cfirst[i] = v[u];
  }
}

// The tile's item table made once per binning (item_tables_kernel) into LDS: its
// SWEEPS * TC entries as 16-byte loads, issued together with the offsets' loads.
template <int BLOCK, int TC>
__device__ __forceinline__ void load_offsets_and_items(const uint32_t *__restrict__ start,
                                                       const uint16_t *__restrict__ table, uint32_t *cfirst,
                                                       uint16_t *items) {
  constexpr int E = (TC + 1 + BLOCK - 1) / BLOCK, Q16 = SWEEPS * TC * 2 / 16, Q = (Q16 + BLOCK - 1) / BLOCK;
  static_assert(SWEEPS * TC * 2 % 16 == 0, "table copy");
  uint32_t v[E];
  uint4 q[Q];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    v[u] = start[i <= TC ? i : TC];
  }
#pragma unroll
  for (int u = 0; u < Q; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    q[u] = reinterpret_cast<const uint4 *>(table)[i < Q16 ? i : Q16 - 1];
  }
#pragma unroll
  for (int u = 0; u < E; ++u) loaded_here(v[u]);
#pragma unroll
  for (int u = 0; u < Q; ++u) {
    loaded_here(q[u].x);
    loaded_here(q[u].y);
    loaded_here(q[u].z);
    loaded_here(q[u].w);
  }
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    if (i <= TC) cfirst[i] = v[u];
  }
#pragma unroll
  for (int u = 0; u < Q; ++u) {
    const int i = threadIdx.x + u * BLOCK;
    if (i < Q16) reinterpret_cast<uint4 *>(items)[i] = q[u];
  }
}

// Item tables of every 3D tile, in the order build_items lists them for workgroups of
// BLOCK work-items: table[tile * SWEEPS * TC + i] and counts[tile].  For bins kept across
// deposits and gathers, which then load the table instead of building it.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) item_tables_kernel(const uint32_t *__restrict__ start,
                                                            uint16_t *__restrict__ table,
                                                            uint32_t *__restrict__ counts) {
  constexpr int TC = TILE_CELLS, PARTS = TC / BLOCK;
  __shared__ uint32_t cfirst[TC + 1];
  __shared__ __align__(16) uint16_t items[SWEEPS * TC];
  __shared__ unsigned long long wave_sums[(BLOCK / 64) * PARTS];
  const int64_t tile = blockIdx.x;
  load_offsets<BLOCK, TC>(start + tile * TC, cfirst);
  __syncthreads();
  const uint32_t total = build_items<BLOCK, TC>(cfirst, items, wave_sums);
  for (uint32_t i = threadIdx.x; i < (uint32_t)(SWEEPS * TC); i += BLOCK)
    table[tile * SWEEPS * TC + i] = i < total ? items[i] : (uint16_t)0;
  if (threadIdx.x == 0) counts[tile] = total;
}

struct Item {
  int64_t begin;  // first marker (binned index)
  uint32_t size;  // markers, at least one
  int cell;       // cell within the tile
};

__device__ __forceinline__ Item item_at(const uint16_t *items, const uint32_t *cfirst, uint32_t it) {
  const uint32_t e = items[it];
  const int c = (int)(e & 0xFFF);
  const uint32_t j = e >> 12;
  const uint32_t first = cfirst[c], count = cfirst[c + 1] - first;
  const uint32_t k = items_of(count);
  uint32_t q, r;
  if (k == 1) {
    q = count;
    r = 0;
  } else if (k == 2) {
    q = count >> 1;
    r = count & 1;
  } else if (k == 3) {
    q = count / 3;
    r = count - 3 * q;
  } else {
    q = count >> 2;
    r = count & 3;
  }
  Item out;
  out.begin = (int64_t)first + j * q + (j < r ? j : r);
  out.size = q + (j < r ? 1u : 0u);
  out.cell = c;
  return out;
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima): every vector-memory load
// of the work-item has completed.
__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// A marker row entry by binned index: with SMALL (every row of the binned arrays ends below
// 4 GiB from its base) the byte offset is a 32-bit value added to the row's base, which
// stays in scalar registers (no 64-bit address arithmetic per access).
template <bool SMALL, class T>
__device__ __forceinline__ T *row_at(T *row, int64_t m) {
  if constexpr (SMALL) return (T *)((char *)row + (uint32_t)((uint32_t)m * (uint32_t)sizeof(T)));
  else return row + m;
}
template <bool SMALL, class T>
__device__ __forceinline__ const T *row_at(const T *row, int64_t m) {
  if constexpr (SMALL) return (const T *)((const char *)row + (uint32_t)((uint32_t)m * (uint32_t)sizeof(T)));
  else return row + m;
}

// Up to ITEM markers of an item, loaded ahead of their use (indices past the item's end
// repeat its first marker, which exists, so every load is in bounds).
template <class T, int C>
struct Ahead {
  T x[ITEM], y[ITEM], z[ITEM], v[ITEM][C > 0 ? C : 1];
};

// The marker rows of a tiled deposit: positions (3) and values (C), each a base pointer.
template <class T, int C>
struct MarkerRows {
  const T *x, *y, *z, *v[C];
};

template <bool SMALL, class T, int C>
__device__ __forceinline__ void load_ahead(const MarkerRows<T, C> &rows, int64_t first, uint32_t count,
                                           Ahead<T, C> &a) {
#pragma unroll
  for (int r = 0; r < ITEM; ++r) {
    const int64_t m = first + (r < (int)count ? r : 0);
    a.x[r] = *row_at<SMALL>(rows.x, m);
    a.y[r] = *row_at<SMALL>(rows.y, m);
    a.z[r] = *row_at<SMALL>(rows.z, m);
#pragma unroll
    for (int q = 0; q < C; ++q) a.v[r][q] = *row_at<SMALL>(rows.v[q], m);
  }
}

// The register sums of one (component, x-slice) pass back to zero (the bias).
template <int S>
__device__ __forceinline__ void rebias(double (&a)[S][S]) {
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int k = 0; k < S; ++k) a[j][k] = BIAS;
}

// The register sums of one pass into the tile's LDS accumulator: node (i, j, k) of the cell
// whose footprint position is `base` sits at base + (i * LY + j) * LZ + k, component q at
// q * L further.  acc - 1.5 * 2^52 as an integer: the low word is the low word of acc's
// bits (the bias has none), the high word one 32-bit subtraction.
template <int S, int LY, int LZ>
__device__ __forceinline__ void flush_slice(unsigned long long *acc, int at, const double (&a)[S][S]) {
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int k = 0; k < S; ++k) atomicAdd(&acc[at + j * LZ + k], (unsigned long long)biased_to_int(a[j][k]));
}

// The deposit of up to ITEM markers of one cell (count >= 1; the first always present):
// their float64 weights once (x weights and y-z products, from the bin's cell), then one
// pass per (component, x-slice) that sums the markers' S^2 contributions exactly in biased
// float64 registers (exact fma, see odd_significand) and adds them to the LDS accumulator.
// A pass holds S^2 registers instead of C S^3.  Slots past `count` hold a copy of the first
// marker with its value scaled by zero: u = odd_significand(0) is the least subnormal, and
// fma(u, w, acc) = acc exactly, so the passes have no per-marker branches.  With ONE, every
// marker is its own pass (for scales under which fewer than ITEM markers fit one register
// sum); the integers are the same either way.
template <int KIND, int P, int C, int S, int L, int LY, int LZ, bool ONE, class T, class Between>
__device__ __forceinline__ void deposit_markers(unsigned long long *acc, int base, const Ahead<T, C> &mk,
                                                uint32_t count, const double (&cell)[3], const AxisRange &ar,
                                                const Grid &g, int num0, int num1, int num2, double scale,
                                                Between &&between) {
  double wx[ITEM][S], wyz[ITEM][S][S], sc[ITEM];
#pragma unroll
  for (int r = 0; r < ITEM; ++r) {
    // the slot's weights from its offsets in the cell, then the full selection under one
    // branch if any axis needs it (a wrap, a clamp or a clamped axis)
    const double rel[3] = {((double)mk.x[r] - g.lower[0]) * g.inverse[0], ((double)mk.y[r] - g.lower[1]) * g.inverse[1],
                           ((double)mk.z[r] - g.lower[2]) * g.inverse[2]};
    const double o[3] = {rel[0] - cell[0], rel[1] - cell[1], rel[2] - cell[2]};
    double w[3][4];
#pragma unroll
    for (int d = 0; d < 3; ++d) offset_weights_f64<KIND, P>(o[d], w[d]);
    if (!offsets_in_cell<KIND, P>(rel, o, ar, g)) {
      cell_weights_f64<KIND, P>(rel[0], cell[0], ar.hi[0], num0, g.code[0], w[0]);
      cell_weights_f64<KIND, P>(rel[1], cell[1], ar.hi[1], num1, g.code[1], w[1]);
      cell_weights_f64<KIND, P>(rel[2], cell[2], ar.hi[2], num2, g.code[2], w[2]);
    }
#pragma unroll
    for (int i = 0; i < S; ++i) wx[r][i] = w[0][i];
#pragma unroll
    for (int j = 0; j < S; ++j)
#pragma unroll
      for (int k = 0; k < S; ++k) wyz[r][j][k] = w[1][j] * w[2][k];
    sc[r] = (r == 0 || (uint32_t)r < count) ? scale : 0.0;
  }
  // LDS reads the caller wants before this deposit's atomics are queued (they complete in
  // order): the weights took long enough for the previous deposit's atomics to drain
  between();
#pragma unroll
  for (int q = 0; q < C; ++q) {
    double vs[ITEM];
#pragma unroll
    for (int r = 0; r < ITEM; ++r) vs[r] = (double)mk.v[r][q] * sc[r];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int at = q * L + base + i * LY * LZ;
      double a[S][S];
#pragma unroll
      for (int r = 0; r < ITEM; ++r) {
        const double u = odd_significand(wx[r][i] * vs[r]);
#pragma unroll
        for (int j = 0; j < S; ++j)
#pragma unroll
          for (int k = 0; k < S; ++k) a[j][k] = __builtin_fma(u, wyz[r][j][k], (r == 0 || ONE) ? BIAS : a[j][k]);
        if (ONE && r + 1 < ITEM) flush_slice<S, LY, LZ>(acc, at, a);
      }
      flush_slice<S, LY, LZ>(acc, at, a);
    }
  }
}

// The field index (times C) of node (gx, gy, z0 - off_z) for each footprint row (ix, iy) of a
// tile, wrapped on periodic axes; -1 for a row beyond a bounded wall (clamp = false) or the
// clamped row (clamp = true, the gather's view).  A row is LZ nodes along z.  The footprint
// of an axis starts OFF nodes before the tile (the uniform stencils' first node), or at the
// tile on a clamped axis, whose extent is num + P - 1 nodes and which no stencil leaves.
template <int LX, int LY, int OFF, int BLOCK>
__device__ __forceinline__ void footprint_rows(const Grid &g, int x0, int y0, int z0, bool clamp, int fc,
                                               long long *rows) {
  const int num0 = (int)g.num[0], num1 = (int)g.num[1];
  const int ext0 = (int)g.extent[0], ext1 = (int)g.extent[1];
  const int o0 = axis_off<OFF>(g.code[0]), o1 = axis_off<OFF>(g.code[1]);
  for (int r = threadIdx.x; r < LX * LY; r += BLOCK) {
    int gx = x0 - o0 + r / LY, gy = y0 - o1 + r % LY;
    bool ok = true;
    if (g.code[0] == 1) gx = gx < 0 ? gx + num0 : (gx >= num0 ? gx - num0 : gx);
    else if (gx < 0 || gx >= ext0) { ok = clamp; gx = gx < 0 ? 0 : ext0 - 1; }
    if (g.code[1] == 1) gy = gy < 0 ? gy + num1 : (gy >= num1 ? gy - num1 : gy);
    else if (gy < 0 || gy >= ext1) { ok = ok && clamp; gy = gy < 0 ? 0 : ext1 - 1; }
    rows[r] = ok ? ((long long)gx * ext1 + gy) * g.extent[2] * fc : -1ll;
  }
}

// The z node of footprint position iz of a tile, wrapped on a periodic axis; -1 beyond a
// bounded wall or a clamped axis's extent (clamp = false) or the clamped node (clamp = true).
__device__ __forceinline__ int footprint_z(const Grid &g, int gz, bool clamp) {
  const int num2 = (int)g.num[2], ext2 = (int)g.extent[2];
  if (g.code[2] == 1) return gz < 0 ? gz + num2 : (gz >= num2 ? gz - num2 : gz);
  if (gz >= 0 && gz < ext2) return gz;
  return clamp ? (gz < 0 ? 0 : ext2 - 1) : -1;
}

// A walk over a tile's footprint, NROWS rows of ROW words (a row: LZ nodes along z times C
// components, consecutive words of the field), in steps of whole rows: work-item t takes
// word k = t % ROW of rows t / ROW, t / ROW + RPS, ... (RPS rows per step), so that its node
// along z, its component and their wrap or clamp are computed once, and an access costs an
// add.  Work-items past RPS * ROW sit out.
template <int BLOCK, int ROW, int NROWS>
struct RowWalk {
  static_assert(ROW <= BLOCK, "a footprint row fits one step");
  static constexpr int RPS = BLOCK / ROW, STEPS = (NROWS + RPS - 1) / RPS;
  int r0, k;
  bool on;
  __device__ explicit RowWalk(int t) : r0(t / ROW), k(t % ROW), on(t < RPS * ROW) {}
  __device__ int row(int u) const { return r0 + u * RPS; }
  __device__ bool has(int u) const { return on && row(u) < NROWS; }
  // a row that exists, for loads issued unconditionally
  __device__ int row_in(int u) const { return has(u) ? row(u) : NROWS - 1; }
};

// The deposit of one tile of TILE_X x TILE_Y x TILE_Z stencil cells.  The bins are in tile
// order, so the tile's cells are consecutive bins; their offsets go to LDS and the markers
// are cut into items (build_items).  A work-item runs over items: it sums an item's
// contributions in exact float64 registers and adds them into the tile's int64 LDS
// accumulator, which spans the tile's footprint before any wrap or clamp.  The markers of
// the next item are loaded while the current one is deposited.  Nodes beyond a bounded wall
// are folded onto the wall node, and the tile is then added into the field once.
// Two waves per SIMD: the item loop needs about 200 registers (capped at 168, for three
// waves, it spills inside the loop and runs at half the speed).
template <class T, int KIND, int P, int C, int BLOCK, bool SMALL, int ZT = 1>
__global__ void __launch_bounds__(BLOCK, BLOCK == 256 ? 2 : 1) scatter_tiled(const T *__restrict__ pos,
                                                                             const T *__restrict__ val,
                                                          int64_t n, const uint32_t *__restrict__ start, Grid g,
                                                          Tiling tl, double *__restrict__ field, int fc,
                                                          const double *__restrict__ slot,
                                                          const uint16_t *__restrict__ tables,
                                                          const uint32_t *__restrict__ table_counts) {
  using Sh = StencilShape<KIND, P>;
  // a workgroup tile of ZT bin tiles along z: consecutive tiles in the bins' key order, so
  // its cells are one contiguous run of bins (bin tile zt's cell k is cell zt * 1024 + k)
  constexpr int TX = TILE_X, TY = TILE_Y, TZ = TILE_Z * ZT, TC = TILE_CELLS * ZT;
  constexpr int S = Sh::S, OFF = Sh::OFF;
  constexpr int LX = TX + S - 1, LY = TY + S - 1, LZ = TZ + S - 1, L = LX * LY * LZ;
  constexpr int PARTS = TC / BLOCK;
  static_assert(TC <= 4096 && TC % BLOCK == 0, "an item holds its cell in 12 bits");
  __shared__ unsigned long long acc[C * L];
  __shared__ uint32_t cfirst[TC + 1];
  __shared__ __align__(16) uint16_t items[SWEEPS * TC];
  __shared__ unsigned long long wave_sums[(BLOCK / 64) * PARTS];
  __shared__ long long rows[LX * LY];

  const int num0 = (int)g.num[0], num1 = (int)g.num[1], num2 = (int)g.num[2];
  const int kx = blockIdx.x % tl.tiles_x;
  const int ky = (blockIdx.x / tl.tiles_x) % tl.tiles_y;
  const int kz = blockIdx.x / (tl.tiles_x * tl.tiles_y);
  const int tx = tile_of_color(kx, tl.cx, tl.color_x);
  const int ty = tile_of_color(ky, tl.cy, tl.color_y);
  const int tz = tile_of_color(kz, tl.cz, tl.color_z);
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int64_t tile_key = (((int64_t)tx * tl.nty + ty) * tl.ntz_bins + (int64_t)tz * ZT) * TILE_CELLS;
  const int t = threadIdx.x;
  // bin tiles of this workgroup tile that exist (the last one along z may hold fewer)
  const int zt_valid = tl.ntz_bins - tz * ZT < ZT ? tl.ntz_bins - tz * ZT : ZT;

  const int64_t tile_id = tile_key / TILE_CELLS;
  // the tile's item count loads with its offsets and items (one memory latency for all)
  const uint32_t tabled_total = (ZT == 1 && tables) ? table_counts[tile_id] : 0u;
  if (ZT == 1 && tables)
    load_offsets_and_items<BLOCK, TC>(start + tile_key, tables + tile_id * SWEEPS * TC, cfirst, items);
  else
    load_offsets_valid<BLOCK, TC>(start + tile_key, zt_valid * TILE_CELLS, cfirst);
  for (int i = t; i < C * L; i += BLOCK) acc[i] = 0ull;
  footprint_rows<LX, LY, OFF, BLOCK>(g, x0, y0, z0, false, fc, rows);
  __syncthreads();
  const uint32_t total =
      (ZT == 1 && tables) ? tabled_total : build_items<BLOCK, TC>(cfirst, items, wave_sums);

  const double scale = uniform_here(slot[0]);
  // markers per register pass (the exact sums stay below 2^51 quanta): below ITEM only for a
  // scale fixed by the caller, when every marker gets its own pass
  const bool one = slot[2] < (double)ITEM;
  // the grid's origin and inverse spacing in scalar registers before the item loop
  Grid gr = g;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    gr.lower[d] = uniform_here(g.lower[d]);
    gr.inverse[d] = uniform_here(g.inverse[d]);
  }

  // The lookahead: the markers of the next item load while the current one is deposited.
  // The loads are unconditional (past the last item they repeat it) and the first ones are
  // drained before the loop, so the wait before a deposit never covers the next item's
  // loads.  The descriptor of the item after next is read from LDS before the current
  // deposit's atomics are queued (LDS operations of a wavefront complete in order).
  MarkerRows<T, C> mrows;
  mrows.x = pos;
  mrows.y = pos + n;
  mrows.z = pos + 2 * n;
#pragma unroll
  for (int q = 0; q < C; ++q) mrows.v[q] = val + q * n;
  const AxisRange ar = axis_ranges(g);
  const uint32_t last = total > 0 ? total - 1 : 0;
  uint32_t it = t;
  Item cur{0, 0, 0}, nxt{0, 0, 0};
  Ahead<T, C> now;
  if (total > 0) {
    cur = item_at(items, cfirst, it < total ? it : last);
    load_ahead<SMALL>(mrows, cur.begin, cur.size, now);
    nxt = item_at(items, cfirst, it + BLOCK < total ? it + BLOCK : last);
    wait_loads();
  }
  while (it < total) {
    Ahead<T, C> next;
    load_ahead<SMALL>(mrows, nxt.begin, nxt.size, next);
    const uint32_t after = it + 2 * BLOCK;
    Item nn;
    const int c = cur.cell, sub = c / TILE_CELLS, loc = c % TILE_CELLS;
    const int cx = loc / (TY * TILE_Z), cy = (loc / TILE_Z) % TY, cz = sub * TILE_Z + loc % TILE_Z;
    const int base = (cx * LY + cy) * LZ + cz;
    const double cell[3] = {(double)(x0 + cx), (double)(y0 + cy), (double)(z0 + cz)};
    const uint32_t first_count = cur.size < (uint32_t)ITEM ? cur.size : ITEM;
    auto read_next = [&]() { nn = item_at(items, cfirst, after < total ? after : last); };
    if (one)
      deposit_markers<KIND, P, C, S, L, LY, LZ, true>(acc, base, now, first_count, cell, ar, gr, num0, num1, num2,
                                                      scale, read_next);
    else
      deposit_markers<KIND, P, C, S, L, LY, LZ, false>(acc, base, now, first_count, cell, ar, gr, num0, num1, num2,
                                                       scale, read_next);
    // items of more than ITEM markers (cells of more than ITEM * SWEEPS markers), ITEM at a time
    for (uint32_t r0 = ITEM; r0 < cur.size; r0 += ITEM) {
      Ahead<T, C> more;
      load_ahead<SMALL>(mrows, cur.begin + r0, cur.size - r0, more);
      const uint32_t k = cur.size - r0 < (uint32_t)ITEM ? cur.size - r0 : ITEM;
      if (one)
        deposit_markers<KIND, P, C, S, L, LY, LZ, true>(acc, base, more, k, cell, ar, gr, num0, num1, num2, scale,
                                                        []() {});
      else
        deposit_markers<KIND, P, C, S, L, LY, LZ, false>(acc, base, more, k, cell, ar, gr, num0, num1, num2, scale,
                                                         []() {});
    }
    cur = nxt;
    nxt = nn;
    now = next;
    it += BLOCK;
  }

  // The tile into the field, one (node, component) per work-item so that neighbouring lanes
  // touch neighbouring words (a footprint row is LZ * C consecutive words of the field, its
  // start in the row table).  Same-colour footprints never overlap, so every node receives
  // at most one addition per launch and the colours run in a fixed order.  Where the
  // footprint's positions are distinct nodes (no periodic axis shorter than the footprint)
  // the addition is a plain read-modify-write whose reads are issued before the barrier
  // that ends the deposit, so they overlap the slowest wavefront's last items; otherwise it
  // is a float64 atomic add performed in L2.  Either way each word becomes old + v * inv
  // rounded once: the same bits whatever the timing.
  const double inv = slot[1];
  const bool plain_z = g.code[2] != 2 && z0 - OFF >= 0 && z0 - OFF + LZ <= num2;  // rows neither wrap nor clamp
  using Walk = RowWalk<BLOCK, LZ * C, LX * LY>;
  const Walk walk(t);
  const int iz = walk.k / C, q = walk.k - iz * C;
  const int gz = plain_z ? z0 - OFF + iz : footprint_z(g, z0 - axis_off<OFF>(g.code[2]) + iz, false);
  const long long zq = (long long)gz * fc + q;
  const bool distinct = (g.code[0] != 1 || num0 >= LX) && (g.code[1] != 1 || num1 >= LY) &&
                        (g.code[2] != 1 || num2 >= LZ);
  double old[Walk::STEPS];
  long long at[Walk::STEPS];
  if (distinct) {
#pragma unroll
    for (int u = 0; u < Walk::STEPS; ++u) {
      const long long rb = rows[walk.row_in(u)];
      at[u] = (walk.has(u) && rb >= 0 && gz >= 0) ? rb + zq : -1ll;
      old[u] = field[at[u] >= 0 ? at[u] : 0];
    }
  }
  __syncthreads();

  // bounded axes: fold the footprint beyond a wall onto the wall node, axis by axis, so
  // that a corner reaches the wall nodes in two steps (the reference folds the stencil)
  const int t0[3] = {x0, y0, z0};
  const int ext[3] = {LX, LY, LZ};
  const int stride[3] = {LY * LZ, LZ, 1};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (g.code[d] != 0) continue;  // periodic: wrapped; clamped: every stencil node exists
    const int num = (int)g.num[d];
    if (t0[d] - OFF >= 0 && t0[d] - OFF + ext[d] - 1 <= num - 1) continue;  // no wall in reach
    for (int i = t; i < L; i += BLOCK) {
      const int id = (i / stride[d]) % ext[d];
      const int gd = t0[d] - OFF + id;
      if (gd >= 0 && gd <= num - 1) continue;
      const int to = gd < 0 ? 0 : num - 1;
      const int j = i + (to - gd) * stride[d];
#pragma unroll
      for (int q = 0; q < C; ++q) atomicAdd(&acc[q * L + j], acc[q * L + i]);
    }
    __syncthreads();
  }

  if (distinct) {
#pragma unroll
    for (int u = 0; u < Walk::STEPS; ++u) {
      const long long v = (long long)acc[q * L + walk.row_in(u) * LZ + iz];
      if (at[u] >= 0 && v != 0) field[at[u]] = old[u] + (double)v * inv;
    }
    return;
  }
  for (int u = 0; u < Walk::STEPS && walk.has(u); ++u) {
    const int r = walk.row(u);
    const long long rb = rows[r];
    const long long v = (long long)acc[q * L + r * LZ + iz];
    if (rb >= 0 && gz >= 0 && v != 0) unsafeAtomicAdd(field + rb + zq, (double)v * inv);
  }
}

// ---- tiled gather ---------------------------------------------------------------------
//
// One workgroup per tile of TILE_X x TILE_Y x TILE_Z stencil cells.  The tile's node
// footprint is staged in LDS as float64 (node-major, the C components of a node together),
// each footprint position holding its node after the periodic wrap or the wall clamp, so
// that stencil node (a, b, e) of a marker is a constant offset from its first node.  The
// tile's markers are one contiguous run of the binned arrays (bins in tile order); the
// workgroup runs over it, one marker per work-item, so every lane has work and consecutive
// lanes read consecutive markers.  Tiles are numbered so that each XCD (workgroups go
// round-robin over the eight) takes a contiguous range of them and neighbouring
// footprints meet in its L2.
template <class T, int KIND, int P, int C, int BLOCK>
__global__ void __launch_bounds__(BLOCK) gather_tiled(const T *__restrict__ pos, int64_t n,
                                                      const uint32_t *__restrict__ start, Grid g, int ntx, int nty,
                                                      int ntz, const double *__restrict__ field, int fc,
                                                      T *__restrict__ res) {
  using Sh = StencilShape<KIND, P>;
  constexpr int TX = TILE_X, TY = TILE_Y, TZ = TILE_Z, TC = TILE_CELLS;
  constexpr int S = Sh::S, OFF = Sh::OFF;
  constexpr int LX = TX + S - 1, LY = TY + S - 1, LZ = TZ + S - 1, L = LX * LY * LZ;
  __shared__ double tile[L * C];

  const int nb = gridDim.x, b = blockIdx.x;
  const int xcd = b % 8, per = nb / 8, extra = nb % 8;
  const int id = xcd * per + (xcd < extra ? xcd : extra) + b / 8;  // the tile, in tile-key order
  const int tz = id % ntz, ty = (id / ntz) % nty, tx = id / (ntz * nty);
  const int num0 = (int)g.num[0], num1 = (int)g.num[1], num2 = (int)g.num[2];
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  // a clamped axis's footprint starts at the tile and runs over num + P - 1 nodes
  const int o0 = axis_off<OFF>(g.code[0]), o1 = axis_off<OFF>(g.code[1]), o2 = axis_off<OFF>(g.code[2]);
#pragma unroll
  for (int i = threadIdx.x; i < L; i += BLOCK) {
    const int iz = i % LZ, iy = (i / LZ) % LY, ix = i / (LZ * LY);
    const int64_t gx = wrap_index(x0 - o0 + ix, g.code[0] == 2 ? g.extent[0] : num0, g.code[0]);
    const int64_t gy = wrap_index(y0 - o1 + iy, g.code[1] == 2 ? g.extent[1] : num1, g.code[1]);
    const int64_t gz = wrap_index(z0 - o2 + iz, g.code[2] == 2 ? g.extent[2] : num2, g.code[2]);
    const int64_t at = ((gx * g.extent[1] + gy) * g.extent[2] + gz) * fc;
#pragma unroll
    for (int c = 0; c < C; ++c) tile[i * C + c] = field[at + c];
  }
  const int64_t tile_key = (int64_t)id * TC;
  const int64_t tile_first = start[tile_key];
  const uint32_t total = start[tile_key + TC] - (uint32_t)tile_first;
  __syncthreads();

  // U markers per work-item at a time: their position loads are in flight together
  constexpr int U = 8;
  for (uint32_t i0 = threadIdx.x; i0 < total; i0 += U * BLOCK) {
    T px[U], py[U], pz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * BLOCK;
      if (i < total) {
        const int64_t m = tile_first + i;
        px[u] = pos[m];
        py[u] = pos[n + m];
        pz[u] = pos[2 * n + m];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * BLOCK;
      if (i >= total) continue;
      const int64_t m = tile_first + i;
      double wx[4], wy[4], wz[4];
      int bx = tile_anchor<KIND, P>(((double)px[u] - g.lower[0]) * g.inverse[0], num0, g.code[0], wx);
      int by = tile_anchor<KIND, P>(((double)py[u] - g.lower[1]) * g.inverse[1], num1, g.code[1], wy);
      int bz = tile_anchor<KIND, P>(((double)pz[u] - g.lower[2]) * g.inverse[2], num2, g.code[2], wz);
      // the anchor cell within the tile (binning put it there; the clamp keeps LDS reads inside)
      bx = (g.code[0] == 1 && bx >= num0 ? bx - num0 : bx) - x0;
      by = (g.code[1] == 1 && by >= num1 ? by - num1 : by) - y0;
      bz = (g.code[2] == 1 && bz >= num2 ? bz - num2 : bz) - z0;
      bx = bx < 0 ? 0 : (bx > TX - 1 ? TX - 1 : bx);
      by = by < 0 ? 0 : (by > TY - 1 ? TY - 1 : by);
      bz = bz < 0 ? 0 : (bz > TZ - 1 ? TZ - 1 : bz);
      const double *node = tile + ((bx * LY + by) * LZ + bz) * C;
      double total_c[C];
#pragma unroll
      for (int c = 0; c < C; ++c) total_c[c] = 0.0;
#pragma unroll
      for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b2 = 0; b2 < S; ++b2) {
          const double wab = wx[a] * wy[b2];
#pragma unroll
          for (int e = 0; e < S; ++e) {
            const double weight = wab * wz[e];
#pragma unroll
            for (int c = 0; c < C; ++c)
              total_c[c] = __builtin_fma(weight, node[((a * LY + b2) * LZ + e) * C + c], total_c[c]);
          }
        }
#pragma unroll
      for (int c = 0; c < C; ++c) res[c * n + m] = (T)total_c[c];
    }
  }
}

// The weights of rel on one axis in float32, from the float64 cell selection of anchor_axis
// (the same cell for every position); returns nothing, the cell comes from the bins.
template <int KIND, int P>
__device__ __forceinline__ void weights_f32(double rel, int num, int code, float *w) {
  const double nd = (double)num;
  double r = rel;
  if (KIND == LINEAR) {
    int l;
    if (code == 1) {
      if (!(r >= 0.0 && r < nd)) r = r - nd * floor(r / nd);
      const int i = (int)fmin(fmax(r, 0.0), nd);
      l = i >= num ? num - 1 : i;
    } else {
      const int i = (int)fmin(fmax(r, 0.0), nd);
      l = i > num - 2 ? num - 2 : i;
    }
    const double d = r - (double)l;
    const float f = (float)(d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d));
    w[0] = 1.0f - f;
    w[1] = f;
    return;
  }
  if (code == 1) {
    if (!(r >= 0.0 && r < nd)) r = r - nd * floor(r / nd);
  } else {
    if (r < 0.0) r = 0.0;
    if (r > nd - 1.0) r = nd - 1.0;
  }
  const double rc = fmin(fmax(r, 0.0), nd);
  const int middle = (P % 2 == 0) ? (int)(rc + 0.5) : (int)rc;
  const float o = (float)(r - (double)middle);
  if (P == 0) {
    w[0] = 1.0f;
  } else if (P == 1) {
    w[0] = 1.0f - o;
    w[1] = o;
  } else if (P == 2) {
    w[0] = 0.5f * (0.5f - o) * (0.5f - o);
    w[1] = 0.75f - o * o;
    w[2] = 0.5f * (0.5f + o) * (0.5f + o);
  } else {
    const float cm = 1.0f - o, sq = o * o, cu = sq * o;
    w[0] = cm * cm * cm * (1.0f / 6.0f);
    w[1] = (3.0f * cu - 6.0f * sq + 4.0f) * (1.0f / 6.0f);
    w[2] = (-3.0f * cu + 3.0f * sq + 3.0f * o + 1.0f) * (1.0f / 6.0f);
    w[3] = cu * (1.0f / 6.0f);
  }
}

// A node of the gather's float32 footprint: one component, or three held as an (x, y)
// pair and a z value (read as one 8-byte and one 4-byte LDS word: no padding word in the
// registers, and no 12-byte read, which the LDS serves at a third of the rate).
struct Node3 {
  float x, y, z;
};
template <int C>
struct NodeSlot {
  using type = Node3;
};
template <>
struct NodeSlot<1> {
  using type = float;
};

// The float32 weights of rel on one axis for a marker binned into stencil cell `cell`
// (in [0, num)).  Where the position needs neither a periodic wrap nor a wall clamp
// (0 <= rel <= hi; rel < num holds on a periodic axis once |o| <= 1/2), the reference's
// offset r - middle is rel - cell exactly: the binning found middle = cell with the same
// float64 arithmetic.  Other positions (a wrap, a clamp, the node past a periodic seam)
// take the full selection of weights_f32.  Linear bounded axes need no test: the left node
// is the bin's cell for any rel and the fraction is clamped.
template <int KIND, int P>
__device__ __forceinline__ void offset_weights_f32(double o, float *w);

template <int KIND, int P>
__device__ __forceinline__ void cell_weights_f32(double rel, double cell, double hi, int num, int code, float *w) {
  if (KIND == SPLINE && code == 2) {  // clamped knots: the float64 Cox-de Boor weights, rounded
    int64_t node[4];
    double wd[4];
    clamped_cell(rel, num, P, 0, node, wd);
#pragma unroll
    for (int a = 0; a <= P; ++a) w[a] = (float)wd[a];
    return;
  }
  const double o = rel - cell;
  if (!offset_in_cell<KIND, P>(rel, o, hi, code)) {
    weights_f32<KIND, P>(rel, num, code, w);
    return;
  }
  offset_weights_f32<KIND, P>(o, w);
}

// The weights of cell_weights_f32 where offset_in_cell holds: float32 polynomials in the
// offset rounded to float32.
template <int KIND, int P>
__device__ __forceinline__ void offset_weights_f32(double o, float *w) {
  if (KIND == LINEAR) {
    const float f = (float)(o < 0.0 ? 0.0 : (o > 1.0 ? 1.0 : o));
    w[0] = 1.0f - f;
    w[1] = f;
    return;
  }
  const float f = (float)o;
  if (P == 0) {
    w[0] = 1.0f;
  } else if (P == 1) {
    w[0] = 1.0f - f;
    w[1] = f;
  } else if (P == 2) {
    w[0] = 0.5f * (0.5f - f) * (0.5f - f);
    w[1] = 0.75f - f * f;
    w[2] = 0.5f * (0.5f + f) * (0.5f + f);
  } else {
    const float cm = 1.0f - f, sq = f * f, cu = sq * f;
    w[0] = cm * cm * cm * (1.0f / 6.0f);
    w[1] = (3.0f * cu - 6.0f * sq + 4.0f) * (1.0f / 6.0f);
    w[2] = (-3.0f * cu + 3.0f * sq + 3.0f * f + 1.0f) * (1.0f / 6.0f);
    w[3] = cu * (1.0f / 6.0f);
  }
}

// Interpolation at one marker of cell (cx, cy, cz) from the (S^3) node slots of the cell,
// in float32: the y-z weight products once, then per x-slice and component a sum over the
// slice's S^2 nodes, then the x weights (S^2 + C S^3 + C S operations).
template <int KIND, int P, int C, int S>
__device__ __forceinline__ void interpolate_marker(const typename NodeSlot<C>::type (&nd)[S][S][S], float x, float y,
                                                   float z, const double (&cell)[3], const AxisRange &ar,
                                                   const Grid &g, int num0, int num1, int num2, float *out) {
  // the weights from the offsets in the cell, then the full selection under one branch if
  // any axis needs it (a wrap, a clamp or a clamped axis)
  float wx[4], wy[4], wz[4];
  const double rel[3] = {((double)x - g.lower[0]) * g.inverse[0], ((double)y - g.lower[1]) * g.inverse[1],
                         ((double)z - g.lower[2]) * g.inverse[2]};
  const double o[3] = {rel[0] - cell[0], rel[1] - cell[1], rel[2] - cell[2]};
  offset_weights_f32<KIND, P>(o[0], wx);
  offset_weights_f32<KIND, P>(o[1], wy);
  offset_weights_f32<KIND, P>(o[2], wz);
  if (!offsets_in_cell<KIND, P>(rel, o, ar, g)) {
    cell_weights_f32<KIND, P>(rel[0], cell[0], ar.hi[0], num0, g.code[0], wx);
    cell_weights_f32<KIND, P>(rel[1], cell[1], ar.hi[1], num1, g.code[1], wy);
    cell_weights_f32<KIND, P>(rel[2], cell[2], ar.hi[2], num2, g.code[2], wz);
  }
  float wyz[S][S];
#pragma unroll
  for (int b = 0; b < S; ++b)
#pragma unroll
    for (int e = 0; e < S; ++e) wyz[b][e] = wy[b] * wz[e];
  float s[C];
#pragma unroll
  for (int a = 0; a < S; ++a) {
    float t[C];
#pragma unroll
    for (int q = 0; q < C; ++q) t[q] = 0.0f;
#pragma unroll
    for (int b = 0; b < S; ++b)
#pragma unroll
      for (int e = 0; e < S; ++e) {
        if constexpr (C == 1) {
          t[0] = fmaf(wyz[b][e], nd[a][b][e], t[0]);
        } else {
          t[0] = fmaf(wyz[b][e], nd[a][b][e].x, t[0]);
          t[1] = fmaf(wyz[b][e], nd[a][b][e].y, t[1]);
          t[2] = fmaf(wyz[b][e], nd[a][b][e].z, t[2]);
        }
      }
#pragma unroll
    for (int q = 0; q < C; ++q) s[q] = a == 0 ? wx[0] * t[q] : fmaf(wx[a], t[q], s[q]);
  }
#pragma unroll
  for (int q = 0; q < C; ++q) out[q] = s[q];
}

// Tiled gather of float32 markers, one marker per work-item.  The tile's node footprint is
// staged in LDS in float32 (each position holding its node after the periodic wrap or the
// wall clamp), the weights are float32 from the float64 cell selection, and the sums are
// float32: the result is float32 either way.  The tile's markers are one contiguous run of
// the binned arrays; consecutive work-items take consecutive markers, so position loads and
// result stores are coalesced, and the lanes of a cell read the same node slots (one LDS
// access serves them).  Each marker's cell comes from a table of cells by marker, filled
// from the cell offsets for chunks of CH markers (a tile holds about 4096 at 4 markers per
// cell).  The next marker's position is loaded while the current one is interpolated.
// Tiles are numbered so that each XCD (workgroups go round-robin over the eight) takes a
// contiguous range of them.
template <int KIND, int P, int C, int BLOCK, bool SMALL, int ZT = 1>
__global__ void __launch_bounds__(BLOCK, 1024 / BLOCK) gather_lanes(const float *__restrict__ pos, int64_t n,
                                                      const uint32_t *__restrict__ start, Grid g, int ntx, int nty,
                                                      int ntz, const double *__restrict__ field, int fc,
                                                      float *__restrict__ res) {
  using Sh = StencilShape<KIND, P>;
  using Slot = typename NodeSlot<C>::type;
  // a workgroup tile of ZT bin tiles along z (consecutive in the bins' key order)
  constexpr int TX = TILE_X, TY = TILE_Y, TZ = TILE_Z * ZT, TC = TILE_CELLS * ZT;
  constexpr int S = Sh::S, OFF = Sh::OFF;
  constexpr int LX = TX + S - 1, LY = TY + S - 1, LZ = TZ + S - 1, L = LX * LY * LZ;
  constexpr int CH = 4096;
  static_assert(C == 1 || C == 3, "float32 gather tiles hold 1 or 3 components");
  static_assert(TC <= 65536, "a cell index fits 16 bits");
  // the footprint: component 0 (C = 1), or the (x, y) pairs and the z values (C = 3)
  __shared__ float2 tile_xy[C == 3 ? L : 1];
  __shared__ float tile_z[L];
  __shared__ uint32_t cfirst[TC + 1];
  __shared__ uint16_t mcell[CH];
  __shared__ long long rows[LX * LY];

  const int nb = gridDim.x, b = blockIdx.x;
  const int xcd = b % 8, per = nb / 8, extra = nb % 8;
  const int id = xcd * per + (xcd < extra ? xcd : extra) + b / 8;  // the tile, in tile-key order
  const int tz = id % ntz, ty = (id / ntz) % nty, tx = id / (ntz * nty);
  const int num0 = (int)g.num[0], num1 = (int)g.num[1], num2 = (int)g.num[2];
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int t = threadIdx.x;
  // ntz counts workgroup tiles; the bins' key layout has ntz_bins 8 x 8 x 16 tiles along z
  const int ntz_bins = (num2 + TILE_Z - 1) / TILE_Z;
  const int64_t tile_key = (((int64_t)tx * nty + ty) * ntz_bins + (int64_t)tz * ZT) * TILE_CELLS;
  const int valid = (ntz_bins - tz * ZT < ZT ? ntz_bins - tz * ZT : ZT) * TILE_CELLS;
  footprint_rows<LX, LY, OFF, BLOCK>(g, x0, y0, z0, true, fc, rows);
  __syncthreads();
  {
    // the cell offsets and the footprint (one (node, component) per work-item and pass; a
    // row is LZ * C consecutive words of the field), every load of the work-item issued
    // before the first one is used: one memory latency for both
    constexpr int EO = (TC + 1 + BLOCK - 1) / BLOCK;
    const bool plain_z = g.code[2] != 2 && z0 - OFF >= 0 && z0 - OFF + LZ <= num2;
    uint32_t o[EO];
#pragma unroll
    for (int u = 0; u < EO; ++u) {
      const int i = t + u * BLOCK;
      o[u] = start[tile_key + (i <= valid ? i : valid)];
    }
    using Walk = RowWalk<BLOCK, LZ * C, LX * LY>;
    const Walk walk(t);
    const int iz = walk.k / C, q = walk.k - iz * C;
    const long long zq =
        (long long)(plain_z ? z0 - OFF + iz : footprint_z(g, z0 - axis_off<OFF>(g.code[2]) + iz, true)) * fc + q;
    double v[Walk::STEPS];
#pragma unroll
    for (int u = 0; u < Walk::STEPS; ++u) v[u] = field[rows[walk.row_in(u)] + zq];  // unconditional loads
#pragma unroll
    for (int u = 0; u < EO; ++u) {
      const int i = t + u * BLOCK;
      if (i <= TC) cfirst[i] = o[u];
    }
    float *flat_xy = reinterpret_cast<float *>(tile_xy);
#pragma unroll
    for (int u = 0; u < Walk::STEPS; ++u) {
      if (walk.has(u)) {
        const int i = walk.row(u) * LZ + iz;
        if (q == C - 1) tile_z[i] = (float)v[u];
        else flat_xy[2 * i + q] = (float)v[u];
      }
    }
  }
  __syncthreads();

  const float *px = pos, *py = pos + n, *pz = pos + 2 * n;
  float *rq[C];
#pragma unroll
  for (int q = 0; q < C; ++q) rq[q] = res + q * n;
  const AxisRange ar = axis_ranges(g);
  const uint32_t first = cfirst[0], count = cfirst[TC] - first;
  for (uint32_t lo = 0; lo < count; lo += CH) {
    const uint32_t hi = count - lo < (uint32_t)CH ? count : lo + CH;
    // the cell of every marker of the chunk [lo, hi)
    for (int c = t; c < TC; c += BLOCK) {
      uint32_t f = cfirst[c] - first, e = cfirst[c + 1] - first;
      f = f > lo ? f : lo;
      e = e < hi ? e : hi;
      for (uint32_t m = f; m < e; ++m) mcell[m - lo] = (uint16_t)c;
    }
    __syncthreads();
    // one marker per work-item; the loads past the chunk repeat its last marker
    const uint32_t mlast = hi - 1;
    uint32_t m = lo + t;
    float ax, ay, az;
    int ac;
    {
      const uint32_t mm = m < hi ? m : mlast;
      const int64_t gm = (int64_t)first + mm;
      ax = *row_at<SMALL>(px, gm);
      ay = *row_at<SMALL>(py, gm);
      az = *row_at<SMALL>(pz, gm);
      ac = mcell[mm - lo];
    }
    for (; m < hi; m += BLOCK) {
      const float x = ax, y = ay, z = az;
      const int c = ac, sub = c / TILE_CELLS, loc = c % TILE_CELLS;
      {
        // the next marker's position and cell, read while this one is interpolated
        const uint32_t mn = m + BLOCK < hi ? m + BLOCK : mlast;
        const int64_t gm = (int64_t)first + mn;
        ax = *row_at<SMALL>(px, gm);
        ay = *row_at<SMALL>(py, gm);
        az = *row_at<SMALL>(pz, gm);
        ac = mcell[mn - lo];
      }
      const int cx = loc / (TY * TILE_Z), cy = (loc / TILE_Z) % TY, cz = sub * TILE_Z + loc % TILE_Z;
      const int base = (cx * LY + cy) * LZ + cz;
      const double cell[3] = {(double)(x0 + cx), (double)(y0 + cy), (double)(z0 + cz)};
      Slot nd[S][S][S];
#pragma unroll
      for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b2 = 0; b2 < S; ++b2)
#pragma unroll
          for (int e = 0; e < S; ++e) {
            const int k = base + (a * LY + b2) * LZ + e;
            if constexpr (C == 1) {
              nd[a][b2][e] = tile_z[k];
            } else {
              const float2 xy = tile_xy[k];
              nd[a][b2][e] = Node3{xy.x, xy.y, tile_z[k]};
            }
          }
      float out[C];
      interpolate_marker<KIND, P, C, S>(nd, x, y, z, cell, ar, g, num0, num1, num2, out);
      const int64_t gm = (int64_t)first + m;
#pragma unroll
      for (int q = 0; q < C; ++q) *row_at<SMALL>(rq[q], gm) = out[q];
    }
    __syncthreads();  // the next chunk rewrites the cell table
  }
}

// The dead bin (markers whose values were all zero at binning) is in no cell; a gather
// still interpolates at those positions, one marker per work-item, its range read from the
// offsets on the device.
template <class T, int KIND, int C, int NDIM = 3>
__global__ void __launch_bounds__(256) gather_dead_kernel(const T *__restrict__ pos, int64_t n, int degree,
                                                          const uint32_t *__restrict__ start, int64_t cells, Grid g,
                                                          const double *__restrict__ field, int fc,
                                                          T *__restrict__ res) {
  const int64_t first = start[cells], last = start[cells + 1];
  for (int64_t m = first + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < last;
       m += (int64_t)gridDim.x * blockDim.x) {
    int64_t node[3][4] = {{0}, {0}, {0}};
    double w[3][4] = {{1.0}, {1.0}, {1.0}};
    int size[3] = {1, 1, 1};
    for (int d = 0; d < NDIM; ++d)
      size[d] = axis_stencil<KIND>(((double)pos[d * n + m] - g.lower[d]) * g.inverse[d], g.num[d], g.code[d], degree,
                                   node[d], w[d]);
    for (int c = 0; c < C; ++c) {
      double total = 0.0;
      for (int a = 0; a < size[0]; ++a)
        for (int b = 0; b < size[1]; ++b) {
          const double wab = NDIM >= 2 ? w[0][a] * w[1][b] : w[0][a];
          for (int e = 0; e < size[2]; ++e) {
            const double weight = NDIM == 3 ? wab * w[2][e] : wab;
            total += weight * field[(node[0][a] * g.stride[0] + node[1][b] * g.stride[1] + node[2][e] * g.stride[2]) * fc + c];
          }
        }
      res[c * n + m] = (T)total;
    }
  }
}

// ---- 2D tiles ----------------------------------------------------------------------------
//
// 2D markers binned in tiles of TILE2 x TILE2 cells.  Deposit: one workgroup per tile, an
// int64 fixed-point accumulator over the tile's footprint (TILE2 + S - 1 nodes per axis) in
// LDS; each work-item walks a contiguous run of the tile's markers (the lanes of a wave on
// different cells) and adds to_fixed(wx * wy * v * 2^k) per node and component -- the
// integers of the one-marker-per-work-item fixed-point deposit (scatter_direct), so the
// result is the same bit for bit; the tile's nonzero words go into the global int64
// accumulator (integer atomics) and a finalize pass converts.  Gather: the tile's float64
// footprint in LDS, one marker per lane, the direct gather's sums in its order.

// The footprint slot of the unwrapped stencil node k of an axis whose footprint starts at
// node `base`: a bounded axis clamps the node (as the stencils do), a periodic one folds.
__device__ __forceinline__ int slot_2d(int k, int base, int num, int code) {
  if (code == 1) {
    const int s = k - base;
    return s < 0 ? s + num : (s >= num ? s - num : s);
  }
  return (k < 0 ? 0 : (k > num - 1 ? num - 1 : k)) - base;
}

template <class T, int KIND, int P, int C>
__global__ void __launch_bounds__(256) scatter_tiled_2d(const T *__restrict__ pos, const T *__restrict__ val,
                                                        int64_t n, const uint32_t *__restrict__ start, Grid g,
                                                        int nty, unsigned long long *__restrict__ accum,
                                                        const double *__restrict__ scalars) {
  using Sh = StencilShape<KIND, P>;
  constexpr int S = Sh::S, OFF = Sh::OFF, L = TILE2 + S - 1, BLOCK = 256;
  __shared__ unsigned long long acc[L * L * C];
  const int tile = blockIdx.x, t = threadIdx.x;
  const int num0 = (int)g.num[0], num1 = (int)g.num[1];
  const int bx = (tile / nty) * TILE2 - OFF, by = (tile % nty) * TILE2 - OFF;
  for (int i = t; i < L * L * C; i += BLOCK) acc[i] = 0ull;
  const uint32_t first = start[(int64_t)tile * TILE2_CELLS], last = start[(int64_t)(tile + 1) * TILE2_CELLS];
  const double scale = scalars[0];
  __syncthreads();
  // runs of RUN consecutive markers (the last clipped at the tile's end), lane t taking run
  // t of each pass of BLOCK runs: a wavefront's row loads cover 64 * RUN consecutive
  // markers (runs a cell apart made every load touch 64 lines: 1024^2, 2^25 markers,
  // quadratic with 3 values 1.86 ms against 0.51 with runs of 8).  A lane sums its run's
  // contributions in int64 registers while the anchor cell stays the same and adds them into
  // the LDS footprint (component-major) when it changes and at the run's end: a few LDS
  // atomics per node and cell instead of one per marker (LDS int64 atomics run at about two
  // lanes per CU clock).  The integer sums are exact, so the field is the direct deposit's
  // bit for bit.  A run is taken four markers at a time, their rows loaded together.
  constexpr uint32_t RUN = 8;
  const uint32_t chunk = RUN;
  constexpr int BATCH = 4;
  for (uint32_t lo = first + t * chunk; lo < last; lo += BLOCK * chunk) {
    const uint32_t hi = lo + chunk < last ? lo + chunk : last;
    int cx = 0, cy = 0;
    bool held = false;
    unsigned long long q[S][S][C];
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
      for (int b = 0; b < S; ++b)
#pragma unroll
        for (int c = 0; c < C; ++c) q[a][b][c] = 0ull;
    auto flush = [&]() {
#pragma unroll
      for (int a = 0; a < S; ++a) {
        const int sx = slot_2d(cx - OFF + a, bx, num0, g.code[0]);
#pragma unroll
        for (int b = 0; b < S; ++b) {
          const int sy = slot_2d(cy - OFF + b, by, num1, g.code[1]);
#pragma unroll
          for (int c = 0; c < C; ++c) {
            atomicAdd(acc + c * (L * L) + sx * L + sy, q[a][b][c]);
            q[a][b][c] = 0ull;
          }
        }
      }
    };
    for (uint32_t m0 = lo; m0 < hi; m0 += BATCH) {
      double px[BATCH], py[BATCH], vs[BATCH][C];
#pragma unroll
      for (int j = 0; j < BATCH; ++j) {
        const uint32_t m = m0 + j < hi ? m0 + j : hi - 1;  // in range: repeats the last marker
        px[j] = (double)pos[m];
        py[j] = (double)pos[n + m];
#pragma unroll
        for (int c = 0; c < C; ++c) vs[j][c] = (double)val[c * n + m] * scale;
      }
#pragma unroll
      for (int j = 0; j < BATCH; ++j) {
        if (m0 + j >= hi) break;
        double wx[4], wy[4];
        const int ax = anchor_axis<KIND, P>((px[j] - g.lower[0]) * g.inverse[0], num0, g.code[0], wx);
        const int ay = anchor_axis<KIND, P>((py[j] - g.lower[1]) * g.inverse[1], num1, g.code[1], wy);
        if (held && (ax != cx || ay != cy)) flush();
        cx = ax;
        cy = ay;
        held = true;
        // value * 2^k: a power-of-two scale, so the integers are round(w * v) * 2^k as in
        // the direct deposit; below 2^50 quanta they take the two-operation rounding
        double vmax = 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c) vmax = fmax(vmax, fabs(vs[j][c]));
        auto accumulate = [&](auto fast) {
#pragma unroll
          for (int a = 0; a < S; ++a)
#pragma unroll
            for (int b = 0; b < S; ++b) {
              const double weight = wx[a] * wy[b];
#pragma unroll
              for (int c = 0; c < C; ++c) {
                const double y = weight * vs[j][c];
                q[a][b][c] += decltype(fast)::value ? to_fixed_small(y) : to_fixed(y);
              }
            }
        };
        if (vmax < 0x1p50)
          accumulate(std::true_type{});
        else
          accumulate(std::false_type{});
      }
    }
    if (held) flush();
  }
  __syncthreads();
  for (int i = t; i < L * L * C; i += BLOCK) {
    const unsigned long long q = acc[i];
    if (q == 0ull) continue;
    const int c = i / (L * L), sl = i - c * (L * L), sx = sl / L, sy = sl % L;
    int gx = bx + sx, gy = by + sy;
    if (g.code[0] == 1) gx = gx < 0 ? gx + num0 : (gx >= num0 ? gx - num0 : gx);
    if (g.code[1] == 1) gy = gy < 0 ? gy + num1 : (gy >= num1 ? gy - num1 : gy);
    atomicAdd(accum + ((int64_t)gx * num1 + gy) * C + c, q);
  }
}

template <class T, int KIND, int P, int C>
__global__ void __launch_bounds__(256) gather_tiled_2d(const T *__restrict__ pos, int64_t n,
                                                       const uint32_t *__restrict__ start, Grid g, int nty,
                                                       const double *__restrict__ field, T *__restrict__ res) {
  using Sh = StencilShape<KIND, P>;
  constexpr int S = Sh::S, OFF = Sh::OFF, L = TILE2 + S - 1, BLOCK = 256;
  __shared__ double tile_f[L * L * C];
  const int tile = blockIdx.x, t = threadIdx.x;
  const int num0 = (int)g.num[0], num1 = (int)g.num[1];
  const int bx = (tile / nty) * TILE2 - OFF, by = (tile % nty) * TILE2 - OFF;
  for (int i = t; i < L * L; i += BLOCK) {
    const int64_t gx = wrap_index(bx + i / L, num0, g.code[0]), gy = wrap_index(by + i % L, num1, g.code[1]);
#pragma unroll
    for (int c = 0; c < C; ++c) tile_f[i * C + c] = field[(gx * num1 + gy) * C + c];
  }
  const uint32_t first = start[(int64_t)tile * TILE2_CELLS], last = start[(int64_t)(tile + 1) * TILE2_CELLS];
  __syncthreads();
  for (uint32_t m = first + t; m < last; m += BLOCK) {
    double wx[4], wy[4];
    const int ax = anchor_axis<KIND, P>(((double)pos[m] - g.lower[0]) * g.inverse[0], num0, g.code[0], wx);
    const int ay = anchor_axis<KIND, P>(((double)pos[n + m] - g.lower[1]) * g.inverse[1], num1, g.code[1], wy);
    double total[C];
#pragma unroll
    for (int c = 0; c < C; ++c) total[c] = 0.0;
#pragma unroll
    for (int a = 0; a < S; ++a) {
      const int sx = slot_2d(ax - OFF + a, bx, num0, g.code[0]);
#pragma unroll
      for (int b = 0; b < S; ++b) {
        const int sy = slot_2d(ay - OFF + b, by, num1, g.code[1]);
        const double weight = wx[a] * wy[b];
#pragma unroll
        for (int c = 0; c < C; ++c) total[c] += weight * tile_f[(sx * L + sy) * C + c];
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) res[c * n + m] = (T)total[c];
  }
}

// ---- derivative gather ------------------------------------------------------------------
//
// The weights and their derivatives on one axis (anchor_axis's cell and weights; the
// derivative of each weight with respect to rel, zero where a bounded axis clamps the
// position or for degree 0), as axis_derivative computes them for the direct kernels.
template <int KIND, int P>
__device__ __forceinline__ int anchor_axis_d(double rel, int num, int code, double *w, double *dw) {
  const int a = anchor_axis<KIND, P>(rel, num, code, w);
  constexpr int S = StencilShape<KIND, P>::S;
  const bool clamped = code == 0 && (rel < 0.0 || rel > (double)(num - 1));
  if (clamped || (KIND == SPLINE && P == 0)) {
#pragma unroll
    for (int i = 0; i < S; ++i) dw[i] = 0.0;
    return a;
  }
  if (KIND == LINEAR || P == 1) {
    dw[0] = -1.0;
    dw[1] = 1.0;
    return a;
  }
  double r = rel;
  if (code == 1 && !(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
  const double o = r - (double)a;  // a is the unwrapped middle node of the wrapped r
  if (P == 2) {
    dw[0] = -(0.5 - o);
    dw[1] = -2.0 * o;
    dw[2] = 0.5 + o;
  } else {
    dw[0] = -0.5 * (1.0 - o) * (1.0 - o);
    dw[1] = 1.5 * o * o - 2.0 * o;
    dw[2] = -1.5 * o * o + o + 0.5;
    dw[3] = 0.5 * o * o;
  }
  return a;
}

// The gradient of the interpolant contracted with per-marker coefficients, for binned 3D
// markers: out[d * os + order[m]] = sum_c coef[c * cs + order[m]] * dF_c/dx_d at binned
// marker m (coefficients and result in the input order).  One workgroup per tile: the tile's
// node footprint in float64 in LDS, float64 weights and sums, U markers per work-item with
// their position loads in flight together.
template <class T, int KIND, int P, int C, int BLOCK>
__global__ void __launch_bounds__(BLOCK) derivative_tiled(const T *__restrict__ pos, int64_t n,
                                                          const uint32_t *__restrict__ start, Grid g, int ntx,
                                                          int nty, int ntz, const double *__restrict__ field,
                                                          const T *__restrict__ coef, int64_t cs,
                                                          const int64_t *__restrict__ order, T *__restrict__ out,
                                                          int64_t os) {
  using Sh = StencilShape<KIND, P>;
  constexpr int TX = TILE_X, TY = TILE_Y, TZ = TILE_Z, TC = TILE_CELLS;
  constexpr int S = Sh::S, OFF = Sh::OFF;
  constexpr int LX = TX + S - 1, LY = TY + S - 1, LZ = TZ + S - 1, L = LX * LY * LZ;
  __shared__ double tile[L * C];

  const int nb = gridDim.x, b = blockIdx.x;
  const int xcd = b % 8, per = nb / 8, extra = nb % 8;
  const int id = xcd * per + (xcd < extra ? xcd : extra) + b / 8;  // the tile, in tile-key order
  const int tz = id % ntz, ty = (id / ntz) % nty, tx = id / (ntz * nty);
  const int num0 = (int)g.num[0], num1 = (int)g.num[1], num2 = (int)g.num[2];
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  for (int i = threadIdx.x; i < L; i += BLOCK) {
    const int iz = i % LZ, iy = (i / LZ) % LY, ix = i / (LZ * LY);
    const int64_t gx = wrap_index(x0 - OFF + ix, num0, g.code[0]);
    const int64_t gy = wrap_index(y0 - OFF + iy, num1, g.code[1]);
    const int64_t gz = wrap_index(z0 - OFF + iz, num2, g.code[2]);
    const int64_t at = ((gx * num1 + gy) * num2 + gz) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) tile[i * C + c] = field[at + c];
  }
  const int64_t tile_key = (int64_t)id * TC;
  const int64_t tile_first = start[tile_key];
  const uint32_t total = start[tile_key + TC] - (uint32_t)tile_first;
  __syncthreads();

  constexpr int U = 4;
  for (uint32_t i0 = threadIdx.x; i0 < total; i0 += U * BLOCK) {
    T px[U], py[U], pz[U];
    int64_t src[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * BLOCK;
      const int64_t m = tile_first + (i < total ? i : 0);
      px[u] = pos[m];
      py[u] = pos[n + m];
      pz[u] = pos[2 * n + m];
      src[u] = order[m];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * BLOCK;
      if (i >= total) continue;
      double wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
      int bx = anchor_axis_d<KIND, P>(((double)px[u] - g.lower[0]) * g.inverse[0], num0, g.code[0], wx, dx);
      int by = anchor_axis_d<KIND, P>(((double)py[u] - g.lower[1]) * g.inverse[1], num1, g.code[1], wy, dy);
      int bz = anchor_axis_d<KIND, P>(((double)pz[u] - g.lower[2]) * g.inverse[2], num2, g.code[2], wz, dz);
      bx = (g.code[0] == 1 && bx >= num0 ? bx - num0 : bx) - x0;
      by = (g.code[1] == 1 && by >= num1 ? by - num1 : by) - y0;
      bz = (g.code[2] == 1 && bz >= num2 ? bz - num2 : bz) - z0;
      bx = bx < 0 ? 0 : (bx > TX - 1 ? TX - 1 : bx);
      by = by < 0 ? 0 : (by > TY - 1 ? TY - 1 : by);
      bz = bz < 0 ? 0 : (bz > TZ - 1 ? TZ - 1 : bz);
      const double *node = tile + ((bx * LY + by) * LZ + bz) * C;
      double k[C];
#pragma unroll
      for (int c = 0; c < C; ++c) k[c] = (double)coef[c * cs + src[u]];
      double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b2 = 0; b2 < S; ++b2) {
          double sw = 0.0, sd = 0.0;
#pragma unroll
          for (int e = 0; e < S; ++e) {
            const double *f = node + ((a * LY + b2) * LZ + e) * C;
            double kf = k[0] * f[0];
#pragma unroll
            for (int c = 1; c < C; ++c) kf = __builtin_fma(k[c], f[c], kf);
            sw = __builtin_fma(wz[e], kf, sw);
            sd = __builtin_fma(dz[e], kf, sd);
          }
          t0 = __builtin_fma(dx[a] * wy[b2], sw, t0);
          t1 = __builtin_fma(wx[a] * dy[b2], sw, t1);
          t2 = __builtin_fma(wx[a] * wy[b2], sd, t2);
        }
      out[src[u]] = (T)(t0 * g.inverse[0]);
      out[os + src[u]] = (T)(t1 * g.inverse[1]);
      out[2 * os + src[u]] = (T)(t2 * g.inverse[2]);
    }
  }
}

// The dead bin's markers (when the bins were made with values) for derivative_tiled, one
// marker per work-item from the global field.
template <class T, int KIND>
__global__ void __launch_bounds__(256) derivative_dead_kernel(const T *__restrict__ pos, int64_t n, int degree,
                                                              const uint32_t *__restrict__ start, int64_t cells, Grid g,
                                                              const double *__restrict__ field, int C,
                                                              const T *__restrict__ coef, int64_t cs,
                                                              const int64_t *__restrict__ order, T *__restrict__ out,
                                                              int64_t os) {
  const int64_t first = start[cells], last = start[cells + 1];
  for (int64_t m = first + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < last;
       m += (int64_t)gridDim.x * blockDim.x) {
    int64_t node[3][4];
    double w[3][4], dw[3][4];
    int size[3];
    for (int d = 0; d < 3; ++d)
      size[d] = axis_derivative<KIND>(((double)pos[d * n + m] - g.lower[d]) * g.inverse[d], g.num[d], g.code[d],
                                      degree, node[d], w[d], dw[d]);
    const int64_t src = order[m];
    double t[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < C; ++c) {
      const double k = (double)coef[c * cs + src];
      for (int a = 0; a < size[0]; ++a)
        for (int b = 0; b < size[1]; ++b)
          for (int e = 0; e < size[2]; ++e) {
            const double f = k * field[((node[0][a] * g.num[1] + node[1][b]) * g.num[2] + node[2][e]) * C + c];
            t[0] += dw[0][a] * w[1][b] * w[2][e] * f;
            t[1] += w[0][a] * dw[1][b] * w[2][e] * f;
            t[2] += w[0][a] * w[1][b] * dw[2][e] * f;
          }
    }
    for (int d = 0; d < 3; ++d) out[d * os + src] = (T)(t[d] * g.inverse[d]);
  }
}

// ---- host side ------------------------------------------------------------------------

struct TileShape {
  int tx, ty, tz, block;
};


static TileShape tile_shape(int kind, int degree, int ncomp) {
  (void)kind;
  (void)degree;
  (void)ncomp;
  return TileShape{TILE_X, TILE_Y, TILE_Z, 256};
}

// Colours along an axis of nt tiles of `tile` cells, `num` cells long, for a stencil of
// `span` nodes: tiles of one colour must have disjoint footprints (nodes they deposit to).
// Neighbouring tiles' footprints overlap, so two colours alternate.  A periodic axis also
// joins tile nt - 1 to tile 0 across the seam and, when the last tile holds at most
// span - 2 cells, tile 0 to tile nt - 2 (tile 0 deposits below node 0, which wraps to the
// last span - 1 nodes and reaches past the short last tile into tile nt - 2's footprint).
// An odd cycle, or that extra join with an even nt, takes three colours: tile 0 alone, then
// tiles 1 .. nt - 1 alternating (tile nt - 1 deposits past the seam onto nodes
// 0 .. span - 2 - OFF only, below the footprint of tile 1).
static int colors_for(int nt, int code, int64_t num, int tile, int span) {
  if (nt == 1) return 1;
  if (code != 1) return 2;
  const int64_t last = num - (int64_t)(nt - 1) * tile;
  return (nt % 2 == 1 || last <= span - 2) ? 3 : 2;
}

// Which (kind, degree, ncomp, dims, codes) the 3D tiled deposit and gather take; others use
// the direct path.  Every degree and component count (groups of three components where a
// kernel holds three, p <= 2; one at a time otherwise); bounded axes and the splines'
// clamped axes of at least a tile, periodic axes of at least two tiles, whole or with a
// partial last tile, and at least a footprint long (a tile's footprint never meets itself
// across the seam); colors_for keeps same-colour footprints apart across the seam, a
// partial last tile included.
static bool tiled_supported(int kind, int degree, int ndim, int64_t ncomp, const Grid &g, const TileShape &ts) {
  if (ndim != 3 || ncomp < 1 || ncomp > 64) return false;
  if (kind == SPLINE && (degree < 0 || degree > 3)) return false;
  const int p = kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree);
  const int span = kind == LINEAR ? 2 : p + 1;  // stencil nodes per axis
  const int t[3] = {ts.tx, ts.ty, ts.tz};
  for (int d = 0; d < 3; ++d) {
    if (g.code[d] == 2 && kind != SPLINE) return false;  // clamped knots belong to the splines
    if (g.num[d] < t[d]) return false;
    if (g.code[d] == 1 && ((g.num[d] + t[d] - 1) / t[d] < 2 || g.num[d] < t[d] + span - 1)) return false;
  }
  return true;
}

// The tiled gradient (derivative_tiled) holds 1 or 3 components up to degree 2, the cubic
// with one, on whole periodic tiles.
static bool derivative_tiled_supported(int kind, int degree, int ndim, int64_t ncomp, const Grid &g,
                                       const TileShape &ts) {
  if (!tiled_supported(kind, degree, ndim, ncomp, g, ts)) return false;
  const int p = kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree);
  if (!((ncomp == 1 || ncomp == 3) && (p <= 2 || (p == 3 && ncomp == 1)))) return false;
  const int t[3] = {ts.tx, ts.ty, ts.tz};
  for (int d = 0; d < 3; ++d)
    if (g.code[d] == 2 || (g.code[d] == 1 && g.num[d] % t[d] != 0)) return false;
  return true;
}

// 2D tiles take every node weighting (linear, quadratic, spline 0..3) and 1 to 3
// components on bounded or periodic axes.
static bool tiled_2d_supported(int kind, int degree, int ndim, int64_t ncomp, const Grid &g) {
  if (ndim != 2 || ncomp < 1 || ncomp > 3) return false;
  if (kind == SPLINE && (degree < 0 || degree > 3)) return false;
  return g.code[0] != 2 && g.code[1] != 2;
}

template <class T, int KIND, int P, int C>
static int run_scatter_2d(const void *pos, const void *val, int64_t n, const uint32_t *start, const Grid &g,
                          double *field, const double *scalars, mpf_context *ctx, hipStream_t stream) {
  const int ntx = (int)((g.num[0] + TILE2 - 1) / TILE2), nty = (int)((g.num[1] + TILE2 - 1) / TILE2);
  const int64_t count = g.num[0] * g.num[1] * C;
  int rc = ensure(ctx->accum, (size_t)count * 8, true, stream);
  if (rc) return rc;
  {
    ProfileScope prof("scatter_tiled_2d", stream);
    hipLaunchKernelGGL((scatter_tiled_2d<T, KIND, P, C>), dim3(ntx * nty), dim3(256), 0, stream, (const T *)pos,
                       (const T *)val, n, start, g, nty, (unsigned long long *)ctx->accum.ptr, scalars);
  }
  MPF_LAUNCH_CHECK();
  return finalize_fixed_with(ctx, field, count, scalars, stream);
}

template <class T, int KIND, int P, int C>
static int run_gather_2d(const void *pos, int64_t n, const uint32_t *start, const Grid &g, const double *field,
                         void *res, hipStream_t stream) {
  const int ntx = (int)((g.num[0] + TILE2 - 1) / TILE2), nty = (int)((g.num[1] + TILE2 - 1) / TILE2);
  {
    ProfileScope prof("gather_tiled_2d", stream);
    hipLaunchKernelGGL((gather_tiled_2d<T, KIND, P, C>), dim3(ntx * nty), dim3(256), 0, stream, (const T *)pos, n,
                       start, g, nty, field, (T *)res);
  }
  MPF_LAUNCH_CHECK();
  const int64_t cells = (int64_t)ntx * nty * TILE2_CELLS;
  hipLaunchKernelGGL((gather_dead_kernel<T, KIND, C, 2>), dim3(64), dim3(256), 0, stream, (const T *)pos, n, P, start,
                     cells, g, field, C, (T *)res);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

// Every 2D case, F(T, KIND, P, C) called with the weighting and component count fixed.
#define CASES_2D_C(F, T, K, P)        \
  if (ncomp == 1) return F(T, K, P, 1); \
  if (ncomp == 2) return F(T, K, P, 2); \
  return F(T, K, P, 3);
#define CASES_2D(F, T)                                          \
  if (kind == LINEAR) { CASES_2D_C(F, T, LINEAR, 1) }             \
  if (kind == QUADRATIC) { CASES_2D_C(F, T, QUADRATIC, 2) }       \
  if (degree == 0) { CASES_2D_C(F, T, SPLINE, 0) }                \
  if (degree == 1) { CASES_2D_C(F, T, SPLINE, 1) }                \
  if (degree == 2) { CASES_2D_C(F, T, SPLINE, 2) }                \
  CASES_2D_C(F, T, SPLINE, 3)

}  // namespace mpf

using namespace mpf;

// The bins a caller keeps: sorted copies of the markers, the permutation and the cell
// offsets.  All device memory owned by the caller; sizes from mpf_bin_sizes.
extern "C" int mpf_bin_sizes(int ndim, const int64_t *num_nodes, int64_t *cells) {
  if (ndim < 1 || ndim > 3) {
    set_error("ndim must be 1, 2 or 3");
    return MPF_ERR_ARGUMENT;
  }
  *cells = bins_of(ndim, num_nodes);  // the offsets array holds bins + 2 entries (the dead bin last)
  return MPF_OK;
}

template <class T, int KIND, int P, int NDIM>
static void launch_histogram(const void *pos, int64_t ps, const void *val, int64_t vs, int64_t n, int ncomp,
                             const Grid &g, uint32_t dead_key, int shift, int groups, int blocks, int64_t chunk,
                             uint32_t *key, uint32_t *hist, unsigned long long *maxbits, hipStream_t stream) {
  hipLaunchKernelGGL((bin_histogram_kernel<T, KIND, P, NDIM>), dim3(blocks), dim3(BIN_THREADS), 0, stream,
                     (const T *)pos, ps, (const T *)val, vs, n, ncomp, g, dead_key, shift, groups, chunk, key, hist,
                     maxbits);
}

template <class T, int KIND, int P>
static void dispatch_histogram_nd(int ndim, const void *pos, int64_t ps, const void *val, int64_t vs, int64_t n,
                                  int ncomp, const Grid &g, uint32_t dead_key, int shift, int groups, int blocks,
                                  int64_t chunk, uint32_t *key, uint32_t *hist, unsigned long long *maxbits,
                                  hipStream_t stream) {
#define HIST(D) \
  launch_histogram<T, KIND, P, D>(pos, ps, val, vs, n, ncomp, g, dead_key, shift, groups, blocks, chunk, key, hist, \
                                  maxbits, stream)
  if (ndim == 1) HIST(1);
  if (ndim == 2) HIST(2);
  if (ndim == 3) HIST(3);
#undef HIST
}

template <class T>
static void dispatch_histogram(int kind, int degree, int ndim, const void *pos, int64_t ps, const void *val,
                               int64_t vs, int64_t n, int ncomp, const Grid &g, uint32_t dead_key, int shift,
                               int groups, int blocks, int64_t chunk, uint32_t *key, uint32_t *hist,
                               unsigned long long *maxbits, hipStream_t stream) {
#define HIST(K, P) \
  dispatch_histogram_nd<T, K, P>(ndim, pos, ps, val, vs, n, ncomp, g, dead_key, shift, groups, blocks, chunk, key, \
                                 hist, maxbits, stream)
  if (kind == LINEAR) HIST(LINEAR, 1);
  else if (kind == QUADRATIC) HIST(QUADRATIC, 2);
  else if (degree == 0) HIST(SPLINE, 0);
  else if (degree == 1) HIST(SPLINE, 1);
  else if (degree == 2) HIST(SPLINE, 2);
  else HIST(SPLINE, 3);
#undef HIST
}

template <class T>
static int launch_group_and_cells(const void *pos, int64_t ps, int ndim, const void *val, int64_t vs, int ncomp,
                                  int64_t n, int shift, int groups, int blocks, int64_t chunk, uint32_t dead_key,
                                  const uint32_t *key, const uint32_t *slots, uint32_t *records, void *spos,
                                  void *sval, int64_t *order, uint32_t *start, hipStream_t stream) {
  if (n > 0) {
#define GROUP(D, C)                                                                                               \
  hipLaunchKernelGGL((bin_group_kernel<T, D, C>), dim3(blocks), dim3(BIN_THREADS), 0, stream, key, (const T *)pos, \
                     ps, ndim, (const T *)val, vs, ncomp, n, shift, groups, chunk, slots, records)
    if (ndim == 3 && ncomp == 1) GROUP(3, 1);
    else if (ndim == 2 && ncomp == 3) GROUP(2, 3);
    else if (ndim == 3 && ncomp == 4) GROUP(3, 4);
    else if (ndim == 3 && ncomp == 3) GROUP(3, 3);
    else if (ndim == 3 && ncomp == 0) GROUP(3, 0);
    else if (ndim == 2 && ncomp == 1) GROUP(2, 1);
    else if (ndim == 2 && ncomp == 2) GROUP(2, 2);
    else if (ndim == 2 && ncomp == 0) GROUP(2, 0);
    else if (ndim == 1 && ncomp == 1) GROUP(1, 1);
    else if (ndim == 1 && ncomp == 0) GROUP(1, 0);
    else GROUP(3, -1);
#undef GROUP
    MPF_LAUNCH_CHECK();
  }
  const dim3 grid(groups), block(BIN_THREADS);
#define STAGED(D, C)                                                                                            \
  if (ndim == D && ncomp == C) {                                                                                \
    hipLaunchKernelGGL((bin_cells_staged<T, D, C>), grid, block, 0, stream, records, n, shift, groups, blocks,    \
                       dead_key, slots, (T *)spos, (T *)sval, order, start);                                    \
    MPF_LAUNCH_CHECK();                                                                                         \
    return MPF_OK;                                                                                              \
  }
  if (shift <= 12) {
    STAGED(3, 1) STAGED(3, 3) STAGED(3, 0) STAGED(2, 1) STAGED(2, 2) STAGED(2, 0) STAGED(1, 1) STAGED(1, 0)
    STAGED(2, 3) STAGED(3, 4)  // path deposits: start binned with (charge, end)
    hipLaunchKernelGGL((bin_cells_direct<T, BIN_KEYS_CAP>), grid, block, 0, stream, records, ndim, ncomp, n, shift,
                       groups, blocks, dead_key, slots, (T *)spos, (T *)sval, order, start);
  } else {
    hipLaunchKernelGGL((bin_cells_direct<T, 32768>), grid, block, 0, stream, records, ndim, ncomp, n, shift, groups,
                       blocks, dead_key, slots, (T *)spos, (T *)sval, order, start);
  }
#undef STAGED
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

// The bins of no markers: every offset 0, and a unit scale (there is nothing to deposit).
__global__ void empty_bins_kernel(uint32_t *__restrict__ start, int64_t count, double *__restrict__ scalars) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    start[i] = 0u;
  if (scalars && blockIdx.x == 0 && threadIdx.x == 0) {
    scalars[0] = 1.0;
    scalars[1] = 1.0;
    scalars[2] = 4.0e9;
  }
}

static int empty_bins(uint32_t *start, int64_t count, double *scalars, hipStream_t stream) {
  const int threads = 256;
  hipLaunchKernelGGL(empty_bins_kernel, dim3(blocks_for(count, threads, 4096)), dim3(threads), 0, stream, start,
                     count, scalars);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

/* Bin markers by stencil cell (tile order in 3D, row-major cells otherwise).  Writes
 * sorted positions (ndim, n) and values (ncomp, n) with contiguous rows, the permutation
 * order[sorted] = original index (int64), the bin offsets start[bins + 2] (uint32;
 * start[bins] is the first dead marker, start[bins + 1] = n) and, in scalars[0..1], the fixed-point scale
 * 2^k, 2^-k and the markers per register pass (scalars[2]); scale_exp overrides k. */
extern "C" int mpf_bin_markers(int kind, int degree, int ndim, const void *positions, int64_t pos_stride,
                               const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp,
                               const double *lower, const double *cell_size, const int64_t *num_nodes,
                               const int64_t *code, void *sorted_positions, void *sorted_values, int64_t *order,
                               uint32_t *start, double *scalars, int scale_exp, mpf_context *ctx, mpf_stream stream_) {
  if (kind < LINEAR || kind > SPLINE || (kind == SPLINE && (degree < 0 || degree > 3))) {
    set_error("bin_markers: kind 0..2 and spline degree 0..3");
    return MPF_ERR_ARGUMENT;
  }
  if (dtype != F32 && dtype != F64) {
    set_error("dtype must be MPF_F32 or MPF_F64, got %d", dtype);
    return MPF_ERR_ARGUMENT;
  }
  if (n < 0 || n > 0x7FFFFFFFll || ncomp < 0 || ncomp > 64 || (ncomp > 0 && !values && n > 0)) {
    set_error("bin_markers: need 0 <= n < 2^31 and 0 <= ncomp <= 64 (0 only without values)");
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size, num_nodes,
                     code);
  if (rc) return rc;
  for (int d = 0; d < ndim; ++d)
    if (code[d] == 2 && (ndim != 3 || kind != SPLINE)) {
      set_error("bin_markers: clamped axes are binned in 3D for the spline kernels only (the direct kernels take "
                "the others)");
      return MPF_ERR_UNSUPPORTED;
    }
  const int64_t cells = bins_of(ndim, num_nodes);  // bins before the dead one
  if (cells + 2 > 0x7FFFFFFFll) {
    set_error("bin_markers: too many cells for 32-bit keys");
    return MPF_ERR_ARGUMENT;
  }
  if (n == 0) return empty_bins(start, cells + 2, scalars, (hipStream_t)stream_);
  // groups of 2^shift keys: about BIN_STAGE / 2 markers per group on average (at most
  // BIN_KEYS_CAP keys), and few enough groups for the LDS counters of passes 1 and 2
  int shift = 0;
  const double keys_per_half_stage = (BIN_STAGE / 2) * (double)cells / (double)(n > 0 ? n : 1);
  while (shift < 12 && (double)(1 << (shift + 1)) <= keys_per_half_stage) ++shift;
  while (shift < 15 && (cells >> shift) + 1 > BIN_GROUPS_CAP) ++shift;
  const int groups = (int)((cells >> shift) + 1);
  if (groups > BIN_GROUPS_CAP) {
    set_error("bin_markers: %lld cells are more than the sort handles", (long long)cells);
    return MPF_ERR_ARGUMENT;
  }
  int64_t blocks = (n + 8191) / 8192;
  if (blocks < 1) blocks = 1;
  if (blocks > BIN_MAX_BLOCKS) blocks = BIN_MAX_BLOCKS;
  const int64_t chunk = n > 0 ? (n + blocks - 1) / blocks : 1;
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  // scratch: the keys in input order, the markers as packed records in group order, the
  // per-(group, block) counts and their scan
  const size_t nn = (size_t)(n > 0 ? n : 1), elem = dtype == F32 ? 4 : 8;
  rc = ensure(ctx->sort_keys, nn * 4, false, stream);
  if (rc) return rc;
  uint32_t *key = (uint32_t *)ctx->sort_keys.ptr;
  rc = ensure(ctx->staging, nn * 4 * (size_t)record_words(ndim, (int)ncomp, (int)elem), false, stream);
  if (rc) return rc;
  uint32_t *records = (uint32_t *)ctx->staging.ptr;
  const size_t entries = (size_t)groups * (size_t)blocks;
  rc = ensure(ctx->offsets, entries * 8, false, stream);
  if (rc) return rc;
  uint32_t *hist = (uint32_t *)ctx->offsets.ptr, *slots = hist + entries;
  rc = ensure(ctx->scalars, 64, true, stream);
  if (rc) return rc;
  unsigned long long *maxbits = (unsigned long long *)ctx->scalars.ptr;
  rc = zero_async(maxbits, 8, stream);
  if (rc) return rc;
  const uint32_t dead_key = (uint32_t)cells;
  if (n > 0) {
    if (dtype == F32)
      dispatch_histogram<float>(kind, degree, ndim, positions, pos_stride, values, val_stride, n, (int)ncomp, g,
                                dead_key, shift, groups, (int)blocks, chunk, key, hist, maxbits, stream);
    else
      dispatch_histogram<double>(kind, degree, ndim, positions, pos_stride, values, val_stride, n, (int)ncomp, g,
                                 dead_key, shift, groups, (int)blocks, chunk, key, hist, maxbits, stream);
    MPF_LAUNCH_CHECK();
  } else {
    rc = zero_async(hist, entries * 4, stream);
    if (rc) return rc;
  }
  size_t temp = 0;
  MPF_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, hist, slots, (int)entries, stream));
  rc = ensure(ctx->sort_tmp, temp, false, stream);
  if (rc) return rc;
  MPF_CHECK(hipcub::DeviceScan::ExclusiveSum(ctx->sort_tmp.ptr, temp, hist, slots, (int)entries, stream));
  if (dtype == F32)
    rc = launch_group_and_cells<float>(positions, pos_stride, ndim, values, val_stride, (int)ncomp, n, shift, groups,
                                       (int)blocks, chunk, dead_key, key, slots, records, sorted_positions,
                                       sorted_values, order, start, stream);
  else
    rc = launch_group_and_cells<double>(positions, pos_stride, ndim, values, val_stride, (int)ncomp, n, shift, groups,
                                        (int)blocks, chunk, dead_key, key, slots, records, sorted_positions,
                                        sorted_values, order, start, stream);
  if (rc) return rc;
  // the scale, from the maximum the count pass found
  rc = prepare_scale(ctx, nullptr, 0, dtype, 1, n, 1.0, scale_exp, stream);
  if (rc) return rc;
  MPF_CHECK(hipMemcpyAsync(scalars, scale_slot(ctx), 24, hipMemcpyDeviceToDevice, stream));
  return MPF_OK;
}

template <class T, int KIND, int P>
static void launch_rebin_count(const void *pos, int64_t ps, const void *val, int64_t vs, int ncomp, const Grid &g,
                               const TileGrid &tg, const uint32_t *ostart, uint32_t *keys, uint2 *movers,
                               uint32_t *moff, uint32_t *total, uint32_t *dead_list, uint32_t *counters,
                               unsigned long long *maxbits, hipStream_t stream) {
  hipLaunchKernelGGL((rebin_count_kernel<T, KIND, P>), dim3(tg.tiles + REBIN_DEAD_GROUPS), dim3(REBIN_THREADS), 0, stream,
                     (const T *)pos, ps, (const T *)val, vs, ncomp, g, tg, ostart, keys, movers, moff, total,
                     dead_list, counters, maxbits);
}

template <class T>
static void dispatch_rebin_count(int kind, int degree, const void *pos, int64_t ps, const void *val, int64_t vs,
                                 int ncomp, const Grid &g, const TileGrid &tg, const uint32_t *ostart, uint32_t *keys,
                                 uint2 *movers, uint32_t *moff, uint32_t *total, uint32_t *dead_list,
                                 uint32_t *counters, unsigned long long *maxbits, hipStream_t stream) {
#define RC(K, P) \
  launch_rebin_count<T, K, P>(pos, ps, val, vs, ncomp, g, tg, ostart, keys, movers, moff, total, dead_list, \
                              counters, maxbits, stream)
  if (kind == LINEAR) RC(LINEAR, 1);
  else if (kind == QUADRATIC) RC(QUADRATIC, 2);
  else if (degree == 0) RC(SPLINE, 0);
  else if (degree == 1) RC(SPLINE, 1);
  else if (degree == 2) RC(SPLINE, 2);
  else RC(SPLINE, 3);
#undef RC
}

static int compose_order(int64_t *order, const int64_t *old_order, int64_t n, hipStream_t stream) {
  if (n == 0) return MPF_OK;
  hipLaunchKernelGGL(compose_order_kernel, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, stream, order, old_order, n);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

static bool ranges_meet(const void *a, size_t na, const void *b, size_t nb) {
  const char *pa = (const char *)a, *pb = (const char *)b;
  return a && b && pa < pb + nb && pb < pa + na;
}

/* Bin again markers that are in the order of earlier bins (old_start, old_order from
 * mpf_bin_markers or this call, same kind/degree/grid), after they moved; the outputs are
 * those of mpf_bin_markers on the same markers (the same bin offsets and scale, each bin's
 * markers in some order), with order[] relative to the markers' first input (old_order
 * composed).  3D: an incremental pass over the tiles (markers that stay in their tile or
 * reach a neighbouring one); it reads back a counter (one stream synchronization) and sorts
 * from scratch when a marker went further.  1D/2D: mpf_bin_markers. */
extern "C" int mpf_rebin_markers(int kind, int degree, int ndim, const void *positions, int64_t pos_stride,
                                 const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp,
                                 const int64_t *old_order, const uint32_t *old_start, const double *lower,
                                 const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                                 void *sorted_positions, void *sorted_values, int64_t *order, uint32_t *start,
                                 double *scalars, int scale_exp, mpf_context *ctx, mpf_stream stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  const size_t elem = dtype == F64 ? 8 : 4;
  if (n > 0 && (ranges_meet(positions, (size_t)(ndim > 0 ? ndim : 0) * pos_stride * elem, sorted_positions,
                            (size_t)(ndim > 0 ? ndim : 0) * n * elem) ||
                ranges_meet(values, (size_t)ncomp * val_stride * elem, sorted_values, (size_t)ncomp * n * elem) ||
                ranges_meet(old_order, (size_t)n * 8, order, (size_t)n * 8) || !old_order || !old_start)) {
    set_error("rebin_markers: the new bins need arrays of their own (they are read from the old ones)");
    return MPF_ERR_ARGUMENT;
  }
  // 3D bins move markers between neighbouring tiles; other bins, and those on clamped axes
  // (cells of their own, clamped_axis_cell), are sorted again
  const bool tiles3d = ndim == 3 && n > 0 && code && code[0] != 2 && code[1] != 2 && code[2] != 2;
  if (!tiles3d) {
    int rc = mpf_bin_markers(kind, degree, ndim, positions, pos_stride, values, val_stride, dtype, n, ncomp, lower,
                             cell_size, num_nodes, code, sorted_positions, sorted_values, order, start, scalars,
                             scale_exp, ctx, stream_);
    if (rc) return rc;
    return compose_order(order, old_order, n, stream);
  }
  if (kind < LINEAR || kind > SPLINE || (kind == SPLINE && (degree < 0 || degree > 3))) {
    set_error("rebin_markers: kind 0..2 and spline degree 0..3");
    return MPF_ERR_ARGUMENT;
  }
  if (dtype != F32 && dtype != F64) {
    set_error("dtype must be MPF_F32 or MPF_F64, got %d", dtype);
    return MPF_ERR_ARGUMENT;
  }
  if (n > 0x7FFFFFFFll || ncomp < 0 || ncomp > 64 || (ncomp > 0 && !values)) {
    set_error("rebin_markers: need 0 <= n < 2^31 and 0 <= ncomp <= 64 (0 only without values)");
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size, num_nodes,
                     code);
  if (rc) return rc;
  TileGrid tg;
  tg.nt[0] = (int)((num_nodes[0] + TILE_X - 1) / TILE_X);
  tg.nt[1] = (int)((num_nodes[1] + TILE_Y - 1) / TILE_Y);
  tg.nt[2] = (int)((num_nodes[2] + TILE_Z - 1) / TILE_Z);
  for (int d = 0; d < 3; ++d) tg.code[d] = (int)code[d];
  const int64_t tiles = (int64_t)tg.nt[0] * tg.nt[1] * tg.nt[2];
  if ((tiles + 1) * TILE_CELLS + 2 > 0x7FFFFFFFll) {
    set_error("rebin_markers: too many cells for 32-bit keys");
    return MPF_ERR_ARGUMENT;
  }
  tg.tiles = (int)tiles;
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  // scratch: keys (input order), movers (the old tiles' ranges), then the dead list, the
  // tile totals, their scan, the movers' offsets by destination and three counters
  const size_t nn = (size_t)n, T1 = (size_t)tiles + 1;
  rc = ensure(ctx->sort_keys, nn * 4, false, stream);
  if (rc) return rc;
  rc = ensure(ctx->staging, nn * 8, false, stream);
  if (rc) return rc;
  const size_t dead_bytes = (nn * 4 + 255) & ~(size_t)255, table_bytes = ((T1 + 1) * 4 + 255) & ~(size_t)255,
               moff_bytes = ((size_t)tiles * REBIN_MOFF * 4 + 255) & ~(size_t)255;
  rc = ensure(ctx->rebin, dead_bytes + 2 * table_bytes + moff_bytes + 256, false, stream);
  if (rc) return rc;
  char *base = (char *)ctx->rebin.ptr;
  uint32_t *dead_list = (uint32_t *)base;
  uint32_t *total = (uint32_t *)(base + dead_bytes), *tstart = (uint32_t *)(base + dead_bytes + table_bytes),
           *moff = (uint32_t *)(base + dead_bytes + 2 * table_bytes),
           *counters = (uint32_t *)(base + dead_bytes + 2 * table_bytes + moff_bytes);
  rc = zero_async(total, table_bytes, stream);
  if (rc) return rc;
  rc = zero_async(counters, 256, stream);
  if (rc) return rc;
  rc = ensure(ctx->scalars, 64, true, stream);
  if (rc) return rc;
  unsigned long long *maxbits = (unsigned long long *)ctx->scalars.ptr;
  rc = zero_async(maxbits, 8, stream);
  if (rc) return rc;
  uint32_t *keys = (uint32_t *)ctx->sort_keys.ptr;
  uint2 *movers = (uint2 *)ctx->staging.ptr;
  if (dtype == F32)
    dispatch_rebin_count<float>(kind, degree, positions, pos_stride, values, val_stride, (int)ncomp, g, tg, old_start,
                                keys, movers, moff, total, dead_list, counters, maxbits, stream);
  else
    dispatch_rebin_count<double>(kind, degree, positions, pos_stride, values, val_stride, (int)ncomp, g, tg,
                                 old_start, keys, movers, moff, total, dead_list, counters, maxbits, stream);
  MPF_LAUNCH_CHECK();
  size_t temp = 0;
  MPF_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, total, tstart, (int)T1, stream));
  rc = ensure(ctx->sort_tmp, temp, false, stream);
  if (rc) return rc;
  MPF_CHECK(hipcub::DeviceScan::ExclusiveSum(ctx->sort_tmp.ptr, temp, total, tstart, (int)T1, stream));
  uint32_t far = 0;
  MPF_CHECK(hipMemcpyAsync(&far, counters + 1, 4, hipMemcpyDeviceToHost, stream));
  MPF_CHECK(hipStreamSynchronize(stream));
  if (far > 0) {  // a marker went further than a neighbouring tile: sort from scratch
    rc = mpf_bin_markers(kind, degree, ndim, positions, pos_stride, values, val_stride, dtype, n, ncomp, lower,
                         cell_size, num_nodes, code, sorted_positions, sorted_values, order, start, scalars,
                         scale_exp, ctx, stream_);
    if (rc) return rc;
    return compose_order(order, old_order, n, stream);
  }
#define PLACE_NC(T, NC)                                                                                               \
  hipLaunchKernelGGL((rebin_place_kernel<T, NC>), dim3(tg.tiles), dim3(REBIN_PLACE_THREADS), 0, stream,                 \
                     (const T *)positions,                                                                            \
                     pos_stride, (const T *)values, val_stride, (int)ncomp, n, tg, old_start, old_order, keys, movers,  \
                     moff, tstart, (T *)sorted_positions, (T *)sorted_values, order, start)
#define PLACE(T)                                                                                                      \
  {                                                                                                                   \
    if (ncomp == 0) PLACE_NC(T, 0);                                                                                   \
    else if (ncomp == 1) PLACE_NC(T, 1);                                                                              \
    else if (ncomp == 3) PLACE_NC(T, 3);                                                                              \
    else PLACE_NC(T, -1);                                                                                             \
    hipLaunchKernelGGL((rebin_dead_kernel<T>), dim3(REBIN_DEAD_GROUPS), dim3(REBIN_THREADS), 0, stream,               \
                       (const T *)positions, pos_stride, (const T *)values, val_stride, (int)ncomp, n, tg, old_start, \
                       old_order, keys, tstart, dead_list, counters, (T *)sorted_positions, (T *)sorted_values,      \
                       order, start);                                                                                 \
  }
  if (dtype == F32) PLACE(float) else PLACE(double)
#undef PLACE
#undef PLACE_NC
  MPF_LAUNCH_CHECK();
  rc = prepare_scale(ctx, nullptr, 0, dtype, 1, n, 1.0, scale_exp, stream);
  if (rc) return rc;
  MPF_CHECK(hipMemcpyAsync(scalars, scale_slot(ctx), 24, hipMemcpyDeviceToDevice, stream));
  return MPF_OK;
}

// ---- launch variants of the 3D tiled deposit and gather ---------------------------------
//
// A variant is a workgroup tile of ZT bin tiles along z (8 x 8 x 16 ZT cells; the bins are
// unchanged) and a workgroup size.  The float32 kernels on 32-bit row offsets come in every
// variant; the table below (mpf_tiled_variant, filled by the autotuner) picks one per
// (operation, kind, degree, components per group); float64 markers take variant 0.
struct Variant {
  int zt, block;
};
constexpr Variant VARIANTS[] = {{1, 256}, {2, 256}, {1, 512}, {2, 512}};
constexpr int NUM_VARIANTS = 4;

static std::mutex g_variant_mutex;
static std::map<std::tuple<int, int, int, int>, int> g_variants;  // (op, kind, degree, C) -> variant

static int variant_for(int op, int kind, int degree, int c) {
  std::lock_guard<std::mutex> lock(g_variant_mutex);
  auto it = g_variants.find({op, kind, kind == SPLINE ? degree : 0, c});
  return it == g_variants.end() ? 0 : it->second;
}

// Whether a workgroup tile of zt bin tiles along z fits this grid: a periodic z axis needs
// two workgroup tiles and a footprint's length, a bounded one a workgroup tile.
static bool variant_fits(int zt, const Grid &g, int span) {
  if (zt == 1) return true;
  const int64_t tz = (int64_t)TILE_Z * zt, nt = (g.num[2] + tz - 1) / tz;
  if (g.code[2] == 1) return nt >= 2 && g.num[2] >= tz + span - 1;
  return g.num[2] >= tz;
}

template <class T, int KIND, int P, int C, int BLOCK, bool SMALL, int ZT>
static int launch_scatter_tiled(const void *pos, const void *val, int64_t n, const uint32_t *start, const Grid &g,
                                double *field, int fc, const double *scalars, const uint16_t *tables,
                                const uint32_t *table_counts, hipStream_t stream) {
  constexpr int TX = TILE_X, TY = TILE_Y, TZ = TILE_Z * ZT;
  Tiling tl;
  tl.tx = TX;
  tl.ty = TY;
  tl.tz = TZ;
  tl.ntx = (int)((g.num[0] + TX - 1) / TX);
  tl.nty = (int)((g.num[1] + TY - 1) / TY);
  tl.ntz = (int)((g.num[2] + TZ - 1) / TZ);
  tl.ntz_bins = (int)((g.num[2] + TILE_Z - 1) / TILE_Z);
  constexpr int SPAN = KIND == LINEAR ? 2 : P + 1;
  tl.cx = colors_for(tl.ntx, g.code[0], g.num[0], TX, SPAN);
  tl.cy = colors_for(tl.nty, g.code[1], g.num[1], TY, SPAN);
  tl.cz = colors_for(tl.ntz, g.code[2], g.num[2], TZ, SPAN);
  auto count_of = [](int nt, int colors, int color) {
    if (colors == 1) return nt;
    if (colors == 3) return color == 2 ? 1 : (nt - color) / 2;  // tiles 1 + color, 3 + color, ...
    return (nt + 1 - color) / 2;
  };
  for (int cx = 0; cx < tl.cx; ++cx)
    for (int cy = 0; cy < tl.cy; ++cy)
      for (int cz = 0; cz < tl.cz; ++cz) {
        tl.color_x = cx;
        tl.color_y = cy;
        tl.color_z = cz;
        tl.tiles_x = count_of(tl.ntx, tl.cx, cx);
        tl.tiles_y = count_of(tl.nty, tl.cy, cy);
        tl.tiles_z = count_of(tl.ntz, tl.cz, cz);
        const int blocks = tl.tiles_x * tl.tiles_y * tl.tiles_z;
        if (blocks == 0) continue;
        ProfileScope prof("scatter_tiled", stream);
        hipLaunchKernelGGL((scatter_tiled<T, KIND, P, C, BLOCK, SMALL, ZT>), dim3(blocks), dim3(BLOCK), 0, stream,
                           (const T *)pos, (const T *)val, n, start, g, tl, field, fc, scalars, tables, table_counts);
        MPF_LAUNCH_CHECK();
      }
  return MPF_OK;
}

template <class T, int KIND, int P, int C>
static int run_scatter_tiled(const void *pos, const void *val, int64_t n, const uint32_t *start, const Grid &g,
                             double *field, int fc, const double *scalars, const uint16_t *tables,
                             const uint32_t *table_counts, hipStream_t stream) {
  // 32-bit row offsets while every row ends below 4 GiB from its base
  const bool small = (uint64_t)n * sizeof(T) < (1ull << 32);
  constexpr int SPAN = KIND == LINEAR ? 2 : P + 1;
  if constexpr (sizeof(T) == 4) {
    if (small) {
      int v = variant_for(0, KIND, P, C);
      if (!variant_fits(VARIANTS[v].zt, g, SPAN)) v = 0;
#define SV(B, Z) \
  launch_scatter_tiled<T, KIND, P, C, B, true, Z>(pos, val, n, start, g, field, fc, scalars, tables, table_counts, stream)
      if (v == 1) return SV(256, 2);
      if (v == 2) return SV(512, 1);
      if (v == 3) return SV(512, 2);
      return SV(256, 1);
#undef SV
    }
  }
  if (small)
    return launch_scatter_tiled<T, KIND, P, C, 256, true, 1>(pos, val, n, start, g, field, fc, scalars, tables,
                                                             table_counts, stream);
  return launch_scatter_tiled<T, KIND, P, C, 256, false, 1>(pos, val, n, start, g, field, fc, scalars, tables,
                                                            table_counts, stream);
}

template <int KIND, int P, int C, int BLOCK, bool SMALL, int ZT>
static void launch_gather_lanes(const void *pos, int64_t n, const uint32_t *start, const Grid &g, const double *field,
                                int fc, void *res, hipStream_t stream) {
  const int ntx = (int)((g.num[0] + TILE_X - 1) / TILE_X), nty = (int)((g.num[1] + TILE_Y - 1) / TILE_Y),
            ntz = (int)((g.num[2] + TILE_Z * ZT - 1) / (TILE_Z * ZT));
  hipLaunchKernelGGL((gather_lanes<KIND, P, C, BLOCK, SMALL, ZT>), dim3(ntx * nty * ntz), dim3(BLOCK), 0, stream,
                     (const float *)pos, n, start, g, ntx, nty, ntz, field, fc, (float *)res);
}

template <class T, int KIND, int P, int C>
static int run_gather_tiled(const void *pos, int64_t n, const uint32_t *start, const Grid &g, const double *field,
                            int fc, void *res, hipStream_t stream) {
  constexpr int TX = TILE_X, TY = TILE_Y, TZ = TILE_Z, BLOCK = 256;
  constexpr int SPAN = KIND == LINEAR ? 2 : P + 1;
  const int ntx = (int)((g.num[0] + TX - 1) / TX), nty = (int)((g.num[1] + TY - 1) / TY),
            ntz = (int)((g.num[2] + TZ - 1) / TZ);
  {
    ProfileScope prof("gather_tiled", stream);
    // 32-bit row offsets while every row ends below 4 GiB from its base
    const bool small = (uint64_t)n * 4u < (1ull << 32);
    if constexpr (sizeof(T) == 4) {
      if (small) {
        int v = variant_for(1, KIND, P, C);
        if (!variant_fits(VARIANTS[v].zt, g, SPAN)) v = 0;
#define GV(B, Z) launch_gather_lanes<KIND, P, C, B, true, Z>(pos, n, start, g, field, fc, res, stream)
        if (v == 1) GV(256, 2);
        else if (v == 2) GV(512, 1);
        else if (v == 3) GV(512, 2);
        else GV(256, 1);
#undef GV
      } else {
        launch_gather_lanes<KIND, P, C, BLOCK, false, 1>(pos, n, start, g, field, fc, res, stream);
      }
    } else {
      hipLaunchKernelGGL((gather_tiled<T, KIND, P, C, BLOCK>), dim3(ntx * nty * ntz), dim3(BLOCK), 0, stream,
                         (const T *)pos, n, start, g, ntx, nty, ntz, field, fc, (T *)res);
    }
  }
  MPF_LAUNCH_CHECK();
  const int64_t cells = (int64_t)ntx * nty * ntz * TILE_CELLS;  // the dead bin follows the tiles' bins
  hipLaunchKernelGGL((gather_dead_kernel<T, KIND, C>), dim3(64), dim3(256), 0, stream, (const T *)pos, n, P, start,
                     cells, g, field, fc, (T *)res);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

// One group of C components (C = 3 for p <= 2, else 1) of a 3D tiled deposit or gather:
// the group's values start at val (rows n apart), its field values at field (fc per node).
template <class T, int C>
static int scatter_tiled_group(int kind, int degree, const void *pos, const void *val, int64_t n,
                               const uint32_t *start, const Grid &g, double *field, int fc, const double *scalars,
                               const uint16_t *tables, const uint32_t *counts, hipStream_t stream) {
#define SC(K, P) run_scatter_tiled<T, K, P, C>(pos, val, n, start, g, field, fc, scalars, tables, counts, stream)
  if (kind == LINEAR) return SC(LINEAR, 1);
  if (kind == QUADRATIC) return SC(QUADRATIC, 2);
  if (degree == 0) return SC(SPLINE, 0);
  if (degree == 1) return SC(SPLINE, 1);
  if (degree == 2) return SC(SPLINE, 2);
  if constexpr (C == 1)
    if (degree == 3) return SC(SPLINE, 3);
#undef SC
  set_error("no tiled kernel for this case");
  return MPF_ERR_UNSUPPORTED;
}

template <class T, int C>
static int gather_tiled_group(int kind, int degree, const void *pos, int64_t n, const uint32_t *start, const Grid &g,
                              const double *field, int fc, void *res, hipStream_t stream) {
#define GA(K, P) run_gather_tiled<T, K, P, C>(pos, n, start, g, field, fc, res, stream)
  if (kind == LINEAR) return GA(LINEAR, 1);
  if (kind == QUADRATIC) return GA(QUADRATIC, 2);
  if (degree == 0) return GA(SPLINE, 0);
  if (degree == 1) return GA(SPLINE, 1);
  if (degree == 2) return GA(SPLINE, 2);
  if constexpr (C == 1)
    if (degree == 3) return GA(SPLINE, 3);
#undef GA
  set_error("no tiled kernel for this case");
  return MPF_ERR_UNSUPPORTED;
}

// The component groups of a 3D tiled call: groups of three components while three remain
// and the kernels hold three (p <= 2), single components otherwise; each group reads the
// same bins.  visit(c0, size) runs one group.
template <class Visit>
static int component_groups(int kind, int degree, int64_t ncomp, Visit &&visit) {
  const int p = kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree);
  for (int64_t c0 = 0; c0 < ncomp;) {
    const int size = (p <= 2 && ncomp - c0 >= 3) ? 3 : 1;
    const int rc = visit(c0, size);
    if (rc) return rc;
    c0 += size;
  }
  return MPF_OK;
}

template <class T>
static int scatter_tiled_3d(int kind, int degree, const void *pos, const void *val, int64_t n, int64_t ncomp,
                            const uint32_t *start, const Grid &g, double *field, const double *scalars,
                            const uint16_t *tables, const uint32_t *counts, hipStream_t stream) {
  return component_groups(kind, degree, ncomp, [&](int64_t c0, int size) {
    const T *v = (const T *)val + c0 * n;
    if (size == 3)
      return scatter_tiled_group<T, 3>(kind, degree, pos, v, n, start, g, field + c0, (int)ncomp, scalars, tables,
                                       counts, stream);
    return scatter_tiled_group<T, 1>(kind, degree, pos, v, n, start, g, field + c0, (int)ncomp, scalars, tables,
                                     counts, stream);
  });
}

template <class T>
static int gather_tiled_3d(int kind, int degree, const void *pos, int64_t n, int64_t ncomp, const uint32_t *start,
                           const Grid &g, const double *field, void *res, hipStream_t stream) {
  return component_groups(kind, degree, ncomp, [&](int64_t c0, int size) {
    T *r = (T *)res + c0 * n;
    if (size == 3)
      return gather_tiled_group<T, 3>(kind, degree, pos, n, start, g, field + c0, (int)ncomp, r, stream);
    return gather_tiled_group<T, 1>(kind, degree, pos, n, start, g, field + c0, (int)ncomp, r, stream);
  });
}

/* Deposit binned markers into field with the tiled kernels (int64 fixed point through
 * float64 exact accumulators; bit-reproducible).  sorted_* and start come from
 * mpf_bin_markers of the same kind/degree/grid; scalars holds [scale, 1/scale].
 * Returns MPF_ERR_UNSUPPORTED for shapes the tiled kernels do not cover. */
extern "C" int mpf_scatter_binned(int kind, int degree, int ndim, const void *sorted_positions,
                                  const void *sorted_values, int dtype, int64_t n, int64_t ncomp,
                                  const uint32_t *start, const uint16_t *item_tables, const uint32_t *item_counts,
                                  const double *scalars, double *field, const double *lower,
                                  const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                                  mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size,
                     num_nodes, code);
  if (rc) return rc;
  const TileShape ts = tile_shape(kind, degree, ncomp);
  if (!tiled_supported(kind, degree, ndim, ncomp, g, ts) && !tiled_2d_supported(kind, degree, ndim, ncomp, g)) {
    set_error("no tiled kernel for kind %d degree %d, %dD, %lld components on this grid", kind, degree, ndim,
              (long long)ncomp);
    return MPF_ERR_UNSUPPORTED;
  }
  if (n == 0) return MPF_OK;  // no markers: nothing to add, no results
  hipStream_t stream = (hipStream_t)stream_;
  if (tiled_2d_supported(kind, degree, ndim, ncomp, g)) {
    if (!ctx) ctx = default_context(stream);
    if (!ctx) return MPF_ERR_HIP;
#define SC2(T, K, P, C) run_scatter_2d<T, K, P, C>(sorted_positions, sorted_values, n, start, g, field, scalars, ctx, stream)
    if (dtype == F32) {
      CASES_2D(SC2, float)
    } else {
      CASES_2D(SC2, double)
    }
#undef SC2
  }
  if (dtype == F32)
    return scatter_tiled_3d<float>(kind, degree, sorted_positions, sorted_values, n, ncomp, start, g, field, scalars,
                                   item_tables, item_counts, stream);
  return scatter_tiled_3d<double>(kind, degree, sorted_positions, sorted_values, n, ncomp, start, g, field, scalars,
                                  item_tables, item_counts, stream);
}

/* Gather at binned markers with the tiled kernels: result (ncomp, n) in the sorted order. */
extern "C" int mpf_gather_binned(int kind, int degree, int ndim, const void *sorted_positions, int dtype, int64_t n,
                                 const uint32_t *start, const uint16_t *item_tables, const uint32_t *item_counts,
                                 const double *field, int64_t ncomp, void *result,
                                 const double *lower, const double *cell_size, const int64_t *num_nodes,
                                 const int64_t *code, mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  (void)item_tables;  // the gather takes one marker per work-item: no work lists
  (void)item_counts;
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size,
                     num_nodes, code);
  if (rc) return rc;
  const TileShape ts = tile_shape(kind, degree, ncomp);
  if (!tiled_supported(kind, degree, ndim, ncomp, g, ts) && !tiled_2d_supported(kind, degree, ndim, ncomp, g)) {
    set_error("no tiled kernel for kind %d degree %d, %dD, %lld components on this grid", kind, degree, ndim,
              (long long)ncomp);
    return MPF_ERR_UNSUPPORTED;
  }
  if (n == 0) return MPF_OK;  // no markers: nothing to add, no results
  hipStream_t stream = (hipStream_t)stream_;
  if (tiled_2d_supported(kind, degree, ndim, ncomp, g)) {
#define GA2(T, K, P, C) run_gather_2d<T, K, P, C>(sorted_positions, n, start, g, field, result, stream)
    if (dtype == F32) {
      CASES_2D(GA2, float)
    } else {
      CASES_2D(GA2, double)
    }
#undef GA2
  }
  if (dtype == F32)
    return gather_tiled_3d<float>(kind, degree, sorted_positions, n, ncomp, start, g, field, result, stream);
  return gather_tiled_3d<double>(kind, degree, sorted_positions, n, ncomp, start, g, field, result, stream);
}

template <class T, int KIND, int P, int C>
static int run_derivative_tiled(const void *pos, int64_t n, const uint32_t *start, const Grid &g, const double *field,
                                const void *coef, int64_t cs, const int64_t *order, void *out, int64_t os,
                                hipStream_t stream) {
  constexpr int BLOCK = 256;
  const int ntx = (int)((g.num[0] + TILE_X - 1) / TILE_X), nty = (int)((g.num[1] + TILE_Y - 1) / TILE_Y),
            ntz = (int)((g.num[2] + TILE_Z - 1) / TILE_Z);
  {
    ProfileScope prof("derivative_tiled", stream);
    hipLaunchKernelGGL((derivative_tiled<T, KIND, P, C, BLOCK>), dim3(ntx * nty * ntz), dim3(BLOCK), 0, stream,
                       (const T *)pos, n, start, g, ntx, nty, ntz, field, (const T *)coef, cs, order, (T *)out, os);
  }
  MPF_LAUNCH_CHECK();
  const int64_t cells = (int64_t)ntx * nty * ntz * TILE_CELLS;
  hipLaunchKernelGGL((derivative_dead_kernel<T, KIND>), dim3(64), dim3(256), 0, stream, (const T *)pos, n, P, start,
                     cells, g, field, C, (const T *)coef, cs, order, (T *)out, os);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

/* The gradient of the interpolant at binned markers, contracted with per-marker
 * coefficients: out[d * out_stride + order[m]] = sum_c coef[c * coef_stride + order[m]] *
 * dF_c/dx_d (float64 sums, the marker dtype out).  sorted_positions, order and start come
 * from mpf_bin_markers of the same kind/degree/grid; coef and out are in the input order. */
extern "C" int mpf_gather_derivative_binned(int kind, int degree, int ndim, const void *sorted_positions, int dtype,
                                            int64_t n, const uint32_t *start, const int64_t *order,
                                            const double *field, int64_t ncomp, const void *coef,
                                            int64_t coef_stride, void *out, int64_t out_stride, const double *lower,
                                            const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                                            mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size,
                     num_nodes, code);
  if (rc) return rc;
  const TileShape ts = tile_shape(kind, degree, ncomp);
  if (!derivative_tiled_supported(kind, degree, ndim, ncomp, g, ts)) {
    set_error("no tiled kernel for kind %d degree %d, %dD, %lld components on this grid", kind, degree, ndim,
              (long long)ncomp);
    return MPF_ERR_UNSUPPORTED;
  }
  if (n == 0) return MPF_OK;  // no markers: nothing to add, no results
  hipStream_t stream = (hipStream_t)stream_;
#define TILED_CASES(X, T)                                                       \
  if (kind == LINEAR && ncomp == 1) return X(T, LINEAR, 1, 1);                  \
  if (kind == LINEAR && ncomp == 3) return X(T, LINEAR, 1, 3);                  \
  if (kind == QUADRATIC && ncomp == 1) return X(T, QUADRATIC, 2, 1);            \
  if (kind == QUADRATIC && ncomp == 3) return X(T, QUADRATIC, 2, 3);            \
  if (kind == SPLINE && degree == 3 && ncomp == 1) return X(T, SPLINE, 3, 1);  \
  if (kind == SPLINE && degree == 2 && ncomp == 1) return X(T, SPLINE, 2, 1);  \
  if (kind == SPLINE && degree == 2 && ncomp == 3) return X(T, SPLINE, 2, 3);  \
  if (kind == SPLINE && degree == 1 && ncomp == 1) return X(T, SPLINE, 1, 1);  \
  if (kind == SPLINE && degree == 1 && ncomp == 3) return X(T, SPLINE, 1, 3);  \
  if (kind == SPLINE && degree == 0 && ncomp == 1) return X(T, SPLINE, 0, 1);  \
  if (kind == SPLINE && degree == 0 && ncomp == 3) return X(T, SPLINE, 0, 3);
#define DV(T, K, P, C) \
  run_derivative_tiled<T, K, P, C>(sorted_positions, n, start, g, field, coef, coef_stride, order, out, out_stride, stream)
  if (dtype == F32) {
    TILED_CASES(DV, float)
  } else {
    TILED_CASES(DV, double)
  }
#undef DV
#undef TILED_CASES
  set_error("no tiled kernel for this case");
  return MPF_ERR_UNSUPPORTED;
}

/* The fixed-point scale of a deposit of these values (rows of ncomp, n markers): scalars
 * gets [2^k, 2^-k, markers per register pass] as mpf_bin_markers writes it, for bins made
 * without values (scale_exp overrides k). */
extern "C" int mpf_fixed_scale(const void *values, int64_t val_stride, int dtype, int64_t n, int64_t ncomp,
                               int scale_exp, double *scalars, mpf_context *ctx, mpf_stream stream_) {
  if (dtype != F32 && dtype != F64) {
    set_error("dtype must be MPF_F32 or MPF_F64, got %d", dtype);
    return MPF_ERR_ARGUMENT;
  }
  if (n < 0 || ncomp < 1 || (!values && n > 0)) {
    set_error("fixed_scale: need n >= 0, ncomp >= 1 and values");
    return MPF_ERR_ARGUMENT;
  }
  hipStream_t stream = (hipStream_t)stream_;
  if (n == 0) return empty_bins(nullptr, 0, scalars, stream);
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  int rc = prepare_scale(ctx, values, val_stride, dtype, ncomp, n, 1.0, scale_exp, stream);
  if (rc) return rc;
  MPF_CHECK(hipMemcpyAsync(scalars, scale_slot(ctx), 24, hipMemcpyDeviceToDevice, stream));
  return MPF_OK;
}

/* The item tables of 3D bins (tile order): tables[tile * 4096 + i] (uint16) and
 * counts[tile], made once so that mpf_scatter_binned / mpf_gather_binned load a tile's
 * work list instead of building it in every launch.  tiles = bins / 1024. */
extern "C" int mpf_bin_items(int ndim, const int64_t *num_nodes, const uint32_t *start, uint16_t *tables,
                             uint32_t *counts, mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  if (ndim != 3) {
    set_error("bin_items: item tables are made for 3D bins");
    return MPF_ERR_UNSUPPORTED;
  }
  const int64_t tiles = bins_of(3, num_nodes) / TILE_CELLS;
  hipStream_t stream = (hipStream_t)stream_;
  if (tiles > 0) {
    hipLaunchKernelGGL(item_tables_kernel<256>, dim3((unsigned)tiles), dim3(256), 0, stream, start, tables, counts);
    MPF_LAUNCH_CHECK();
  }
  return MPF_OK;
}

// out[c, order[m]] = sorted[c, m]: the reads coalesced, the writes scattered by the
// permutation.  Up to 512 MB of rows one work-item per marker writes all its rows (its
// input index read once); above, one launch per row keeps each wavefront's scattered
// writes within one row (3 rows of 2^26 float32: 10.0 ms row by row against 13.8 for all
// rows at once; 3 rows of 2^25: 5.2 against 5.6 the other way round).
template <class T>
__global__ void __launch_bounds__(256) unsort_rows_kernel(const T *__restrict__ sorted, T *__restrict__ out,
                                                          int64_t out_stride, const int64_t *__restrict__ order,
                                                          int64_t rows, int64_t n) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = order[m];
    for (int64_t c = 0; c < rows; ++c) out[c * out_stride + o] = sorted[c * n + m];
  }
}

template <class T>
__global__ void __launch_bounds__(256) unsort_row_kernel(const T *__restrict__ sorted, T *__restrict__ out,
                                                         const int64_t *__restrict__ order, int64_t n) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x)
    out[order[m]] = sorted[m];
}

namespace mpf {

// Rows (rows, n) in bins order, contiguous, back to the input order into rows out_stride
// elements apart.
static int unsort_rows_strided(const void *sorted, void *out, int64_t out_stride, const int64_t *order, int64_t rows,
                               int64_t n, int dtype, hipStream_t stream) {
  if (rows == 0 || n == 0) return MPF_OK;
  const int blocks = blocks_for(n, 256, 1 << 16);
  const int64_t bytes = rows * n * (dtype == F32 ? 4 : 8);
  if (bytes <= (int64_t)512 << 20) {
    if (dtype == F32)
      hipLaunchKernelGGL(unsort_rows_kernel<float>, dim3(blocks), dim3(256), 0, stream, (const float *)sorted,
                         (float *)out, out_stride, order, rows, n);
    else
      hipLaunchKernelGGL(unsort_rows_kernel<double>, dim3(blocks), dim3(256), 0, stream, (const double *)sorted,
                         (double *)out, out_stride, order, rows, n);
    MPF_LAUNCH_CHECK();
    return MPF_OK;
  }
  for (int64_t c = 0; c < rows; ++c) {
    if (dtype == F32)
      hipLaunchKernelGGL(unsort_row_kernel<float>, dim3(blocks), dim3(256), 0, stream, (const float *)sorted + c * n,
                         (float *)out + c * out_stride, order, n);
    else
      hipLaunchKernelGGL(unsort_row_kernel<double>, dim3(blocks), dim3(256), 0, stream,
                         (const double *)sorted + c * n, (double *)out + c * out_stride, order, n);
    MPF_LAUNCH_CHECK();
  }
  return MPF_OK;
}

static std::atomic<int64_t> g_binned_min{BINNED_MIN_MARKERS};

int64_t binned_min() { return g_binned_min.load(); }

bool binned_route(int kind, int degree, int ndim, int64_t ncomp, const Grid &g, int64_t n) {
  if (n < binned_min() || n > 0x7FFFFFFFll) return false;  // 32-bit bin offsets
  return tiled_supported(kind, degree, ndim, ncomp, g, tile_shape(kind, degree, ncomp)) ||
         tiled_2d_supported(kind, degree, ndim, ncomp, g);
}

// The bins of one call, carved from ctx->bins: sorted positions and values, results in bins
// order, the permutation, the bin offsets and the scale, each on a 256-byte boundary.
struct CallBins {
  void *pos, *val, *res;
  int64_t *order;
  uint32_t *start;
  double *scalars;
};

static int call_bins(mpf_context *ctx, int ndim, const int64_t *num, int64_t n, int64_t val_rows, int64_t res_rows,
                     size_t elem, CallBins &b, hipStream_t stream) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t cells = (size_t)bins_of(ndim, num);
  const size_t pos_b = up((size_t)ndim * n * elem), val_b = up((size_t)val_rows * n * elem),
               res_b = up((size_t)res_rows * n * elem), ord_b = up((size_t)n * 8), st_b = up((cells + 2) * 4);
  int rc = ensure(ctx->bins, pos_b + val_b + res_b + ord_b + st_b + 256, false, stream);
  if (rc) return rc;
  char *p = (char *)ctx->bins.ptr;
  b.pos = p;
  b.val = p + pos_b;
  b.res = p + pos_b + val_b;
  b.order = (int64_t *)(p + pos_b + val_b + res_b);
  b.start = (uint32_t *)(p + pos_b + val_b + res_b + ord_b);
  b.scalars = (double *)(p + pos_b + val_b + res_b + ord_b + st_b);
  return MPF_OK;
}

int scatter_via_bins(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, const void *values,
                     int64_t val_stride, int dtype, int64_t n, int64_t ncomp, double *field, const double *lower,
                     const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                     hipStream_t stream) {
  CallBins b;
  int rc = call_bins(ctx, ndim, num_nodes, n, ncomp, 0, dtype == F32 ? 4 : 8, b, stream);
  if (rc) return rc;
  rc = mpf_bin_markers(kind, degree, ndim, positions, pos_stride, values, val_stride, dtype, n, ncomp, lower,
                       cell_size, num_nodes, code, b.pos, b.val, b.order, b.start, b.scalars, MPF_SCALE_AUTO, ctx,
                       stream);
  if (rc) return rc;
  return mpf_scatter_binned(kind, degree, ndim, b.pos, b.val, dtype, n, ncomp, b.start, nullptr, nullptr, b.scalars,
                            field, lower, cell_size, num_nodes, code, ctx, stream);
}

int gather_via_bins(int kind, int degree, int ndim, const void *positions, int64_t pos_stride, int dtype, int64_t n,
                    const double *field, int64_t ncomp, void *result, int64_t result_stride, const double *lower,
                    const double *cell_size, const int64_t *num_nodes, const int64_t *code, mpf_context *ctx,
                    hipStream_t stream) {
  CallBins b;
  int rc = call_bins(ctx, ndim, num_nodes, n, 0, ncomp, dtype == F32 ? 4 : 8, b, stream);
  if (rc) return rc;
  rc = mpf_bin_markers(kind, degree, ndim, positions, pos_stride, nullptr, 0, dtype, n, 0, lower, cell_size,
                       num_nodes, code, b.pos, nullptr, b.order, b.start, b.scalars, MPF_SCALE_AUTO, ctx, stream);
  if (rc) return rc;
  rc = mpf_gather_binned(kind, degree, ndim, b.pos, dtype, n, b.start, nullptr, nullptr, field, ncomp, b.res, lower,
                         cell_size, num_nodes, code, ctx, stream);
  if (rc) return rc;
  if (ncomp >= 2 && n >= PERM_MIN && perm_plan_ok(n))  // one row: scattered writes cost about the plan
    return unsort_planned(b.res, result, result_stride, b.order, ncomp, n, dtype, ctx, stream);
  return unsort_rows_strided(b.res, result, result_stride, b.order, ncomp, n, dtype, stream);
}

}  // namespace mpf

/* The launch variant of the float32 3D tiled deposit (op 0) or gather (op 1) for a (kind,
 * degree, components per group): returns the previous one, variant < 0 only queries. */
extern "C" int mpf_tiled_variant(int op, int kind, int degree, int ncomp_group, int variant) {
  if (op < 0 || op > 1 || kind < LINEAR || kind > SPLINE || (ncomp_group != 1 && ncomp_group != 3) ||
      variant >= NUM_VARIANTS) {
    set_error("tiled_variant: op 0/1, kind 0..2, 1 or 3 components per group, variant below %d", NUM_VARIANTS);
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_variant_mutex);
  const auto key = std::make_tuple(op, kind, kind == SPLINE ? degree : 0, ncomp_group);
  auto it = g_variants.find(key);
  const int old = it == g_variants.end() ? 0 : it->second;
  if (variant >= 0) g_variants[key] = variant;
  return old;
}

/* The shape of launch variant `variant`: tile[3] cells per workgroup tile and block[0] work-items. */
extern "C" int mpf_tiled_variant_shape(int variant, int *tile, int *block) {
  if (variant < 0 || variant >= NUM_VARIANTS) {
    set_error("tiled_variant_shape: variants are 0..%d", NUM_VARIANTS - 1);
    return MPF_ERR_ARGUMENT;
  }
  tile[0] = TILE_X;
  tile[1] = TILE_Y;
  tile[2] = TILE_Z * VARIANTS[variant].zt;
  *block = VARIANTS[variant].block;
  return MPF_OK;
}

/* The smallest batch that mpf_scatter / mpf_gather bin inside the call; returns the previous
 * value (a negative argument only queries). */
extern "C" int64_t mpf_set_binned_min(int64_t markers) {
  if (markers < 0) return g_binned_min.load();
  return g_binned_min.exchange(markers);
}

/* Rows (rows, n) in bins order back to the input order: out[c * n + order[m]] =
 * sorted[c * n + m], order a permutation of 0..n-1 (mpf_bin_markers'). */
extern "C" int mpf_unsort_rows(const void *sorted, void *out, const int64_t *order, int64_t rows, int64_t n,
                               int dtype, mpf_stream stream_) {
  if (dtype != F32 && dtype != F64) {
    set_error("dtype must be MPF_F32 or MPF_F64, got %d", dtype);
    return MPF_ERR_ARGUMENT;
  }
  if (rows < 0 || n < 0) {
    set_error("unsort_rows: need rows >= 0 and n >= 0");
    return MPF_ERR_ARGUMENT;
  }
  hipStream_t stream = (hipStream_t)stream_;
  if (rows >= 2 && n >= PERM_MIN && perm_plan_ok(n)) {  // a plan in the stream's default context
    mpf_context *ctx = default_context(stream);
    if (!ctx) return MPF_ERR_HIP;
    return unsort_planned(sorted, out, n, order, rows, n, dtype, ctx, stream);
  }
  return unsort_rows_strided(sorted, out, n, order, rows, n, dtype, stream);
}
