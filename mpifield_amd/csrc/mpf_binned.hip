// mpf_binned.hip -- the hot path: markers binned by stencil cell with an on-device
// counting sort, then deposited and gathered one tile of cells per workgroup.
//
// Binning.  A marker's stencil cell ("anchor") is the cell whose nodes it touches, found
// with the same float64 arithmetic as the reference's stencils (quadratic:
// int(r + 1/2); linear: the left node; spline p: int(r + 1/2) for even p, int(r) for odd p;
// wrapped on a periodic axis, clamped on a bounded one).  Keys are row-major anchor
// indices; a count pass (atomic ranks), an exclusive scan and a placement pass give the
// markers sorted by anchor, the cell offsets and the permutation.  Markers whose values
// are all zero (dead) go to a trailing bin that deposits nothing.
//
// Deposit.  One workgroup per tile of TX x TY x TZ anchor cells; one lane per cell.  A
// lane runs over its cell's markers and accumulates the (p+1)^3 x C contributions in
// float64 registers as exact integers: every contribution is rounded to an integer
// number of fixed-point quanta inside one fma on an accumulator biased by 1.5 * 2^52, where
// the unit in the last place is one quantum; the factors are arranged so that the exact
// product is never a tie (see exact_fma below), hence the sum is independent of the
// marker order.  Each lane then adds its integers into the tile's int64 LDS accumulator
// (covering the tile plus its stencil halo), and the workgroup adds the tile into the
// float64 field once.  Tiles whose footprints overlap run in different launches (a 2- or
// 3-colouring per axis), in a fixed order, so the field update needs no atomics and is
// bit-reproducible.
//
// Gather.  One workgroup per tile: the tile's node footprint is staged in LDS; a lane
// loads its cell's (p+1)^3 x C node values into registers once and interpolates each of
// its markers, with the weight products and the sum in the reference's order
// (gather_quadratic_3d_pyccel.py:103-114).
#include <hipcub/hipcub.hpp>

#include "mpf_internal.h"

namespace mpf {

// ---- per-axis stencil cell and weights -----------------------------------------------

// The anchor (unwrapped) and the weights of rel on one axis; nodes are anchor - OFF + a,
// a = 0..P, OFF = P / 2 (linear: 0), clamped to [0, num - 1] on a bounded axis and taken
// modulo num on a periodic one.  Same arithmetic as mpf_device.h's stencils.
template <int KIND, int P>
__device__ __forceinline__ int64_t anchor_axis(double rel, int64_t num, int code, double *w) {
  double r = rel;
  if (KIND == LINEAR) {
    int64_t l;
    if (code == 1) {
      r = periodic_position(r, num);
      const int64_t i = (int64_t)r;
      l = i >= num ? num - 1 : i;
    } else {
      const int64_t i = r >= 0.0 ? (int64_t)r : 0;
      l = i > num - 2 ? num - 2 : i;
    }
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
  const int64_t middle = (P % 2 == 0) ? (int64_t)(r + 0.5) : (int64_t)r;
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

template <class T, int KIND, int P, int NDIM>
__global__ void __launch_bounds__(256) bin_count_kernel(const T *__restrict__ pos, int64_t ps,
                                                        const T *__restrict__ val, int64_t vs, int64_t n, int ncomp,
                                                        Grid g, int64_t dead_key, uint32_t *__restrict__ key,
                                                        uint32_t *__restrict__ rank, uint32_t *__restrict__ count,
                                                        unsigned long long *__restrict__ maxbits) {
  double vmax = 0.0;
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    bool live = ncomp == 0;  // a gather's markers carry no values and are all live
    for (int c = 0; c < ncomp; ++c) {
      const double v = fabs((double)val[c * vs + m]);
      live |= v != 0.0;
      if (v > vmax && v < 1.0e300) vmax = v;
    }
    int64_t k = dead_key;
    if (live) {
      k = 0;
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        double w[4];
        const double rel = ((double)pos[d * ps + m] - g.lower[d]) * g.inverse[d];
        int64_t a = anchor_axis<KIND, P>(rel, g.num[d], g.code[d], w);
        if (g.code[d] == 1 && a >= g.num[d]) a -= g.num[d];
        k = k * g.num[d] + a;
      }
    }
    key[m] = (uint32_t)k;
    rank[m] = atomicAdd(count + k, 1u);
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(vmax, off, 64);
    vmax = o > vmax ? o : vmax;
  }
  if ((threadIdx.x & 63) == 0 && vmax > 0.0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(vmax));
}

template <class T>
__global__ void __launch_bounds__(256) bin_place_kernel(const T *__restrict__ pos, int64_t ps, int ndim,
                                                        const T *__restrict__ val, int64_t vs, int ncomp, int64_t n,
                                                        const uint32_t *__restrict__ key,
                                                        const uint32_t *__restrict__ rank,
                                                        const uint32_t *__restrict__ start, T *__restrict__ spos,
                                                        T *__restrict__ sval, int64_t *__restrict__ order) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    const int64_t dst = (int64_t)start[key[m]] + rank[m];
    for (int d = 0; d < ndim; ++d) spos[d * n + dst] = pos[d * ps + m];
    for (int c = 0; c < ncomp; ++c) sval[c * n + dst] = val[c * vs + m];
    order[dst] = m;
  }
}

// ---- exact fixed-point accumulation in float64 registers --------------------------------
//
// acc holds 1.5 * 2^52 + s for an integer s with |s| < 2^51, so its last place is 1.
// exact_fma(u, w, acc) = fma(u, w, acc) rounds acc + u*w to the nearest integer once.  With
// u's lowest significand bit forced to one and |u| < 2^51, u = odd * 2^e with e <= -2, and
// w in [0, 1] is odd * 2^f with f <= 0: the exact product is an odd multiple of a power of
// two at most 1/4, so it is never halfway between two integers and the rounding is the
// same whatever acc holds.  The accumulated integer is therefore the sum of the
// contributions each rounded to the nearest quantum, in any order.
__device__ __forceinline__ double odd_significand(double u) {
  long long b = __double_as_longlong(u);
  return __longlong_as_double(b | 1ll);
}

__device__ __forceinline__ long long biased_to_int(double acc) {
  return __double_as_longlong(acc) - 0x4338000000000000ll;  // acc - 1.5 * 2^52, as an integer
}

constexpr double BIAS = 6755399441055744.0;  // 1.5 * 2^52

// ---- tiled deposit --------------------------------------------------------------------

struct Tiling {
  int tx, ty, tz;          // tile extent in anchor cells
  int ntx, nty, ntz;       // tiles per axis
  int cx, cy, cz;          // colours per axis
  int color_x, color_y, color_z;  // the colour of this launch
  int tiles_x, tiles_y, tiles_z;  // tiles of this colour per axis
};

__device__ __forceinline__ int color_of(int t, int nt, int colors) {
  if (colors == 2) return t & 1;
  if (colors == 3) return t == nt - 1 ? 2 : (t & 1);
  return 0;
}

// The k-th tile of colour `color` along an axis of nt tiles.
__device__ __forceinline__ int tile_of_color(int k, int nt, int colors, int color) {
  if (colors == 1) return k;
  if (colors == 3 && color == 2) return nt - 1;
  return 2 * k + color;
}

template <class T, int KIND, int P, int C, int TX, int TY, int TZ, int BLOCK>
__global__ void __launch_bounds__(BLOCK) scatter_tiled(const T *__restrict__ pos, const T *__restrict__ val,
                                                       int64_t n, const uint32_t *__restrict__ start, Grid g,
                                                       Tiling tl, double *__restrict__ field,
                                                       const double *__restrict__ slot) {
  using Sh = StencilShape<KIND, P>;
  constexpr int S = Sh::S, OFF = Sh::OFF;
  constexpr int LX = TX + S - 1, LY = TY + S - 1, LZ = TZ + S - 1, L = LX * LY * LZ;
  __shared__ unsigned long long acc[C * L];

  const int kx = blockIdx.x % tl.tiles_x;
  const int ky = (blockIdx.x / tl.tiles_x) % tl.tiles_y;
  const int kz = blockIdx.x / (tl.tiles_x * tl.tiles_y);
  const int tx = tile_of_color(kx, tl.ntx, tl.cx, tl.color_x);
  const int ty = tile_of_color(ky, tl.nty, tl.cy, tl.color_y);
  const int tz = tile_of_color(kz, tl.ntz, tl.cz, tl.color_z);
  const int64_t x0 = (int64_t)tx * TX, y0 = (int64_t)ty * TY, z0 = (int64_t)tz * TZ;

  for (int i = threadIdx.x; i < C * L; i += BLOCK) acc[i] = 0ull;
  __syncthreads();

  const double scale = slot[0], inv = slot[1];
  const int64_t chunk = (int64_t)slot[2];
  for (int cl = threadIdx.x; cl < TX * TY * TZ; cl += BLOCK) {
    const int lz = cl % TZ, ly = (cl / TZ) % TY, lx = cl / (TZ * TY);
    const int64_t ax = x0 + lx, ay = y0 + ly, az = z0 + lz;
    const bool inside = ax < g.num[0] && ay < g.num[1] && az < g.num[2];
    const int64_t cell = (ax * g.num[1] + ay) * g.num[2] + az;
    const int64_t first = inside ? (int64_t)start[cell] : 0;
    const int64_t last = inside ? (int64_t)start[cell + 1] : 0;
    double a[C][S][S][S];
    int64_t done = first;
    while (done < last) {
      const int64_t stop = last - done > chunk ? done + chunk : last;
#pragma unroll
      for (int c = 0; c < C; ++c)
#pragma unroll
        for (int i = 0; i < S; ++i)
#pragma unroll
          for (int j = 0; j < S; ++j)
#pragma unroll
            for (int k = 0; k < S; ++k) a[c][i][j][k] = BIAS;
      for (int64_t m = done; m < stop; ++m) {
        double wx[4], wy[4], wz[4];
        anchor_axis<KIND, P>(((double)pos[m] - g.lower[0]) * g.inverse[0], g.num[0], g.code[0], wx);
        anchor_axis<KIND, P>(((double)pos[n + m] - g.lower[1]) * g.inverse[1], g.num[1], g.code[1], wy);
        anchor_axis<KIND, P>(((double)pos[2 * n + m] - g.lower[2]) * g.inverse[2], g.num[2], g.code[2], wz);
        double wyz[S][S];
#pragma unroll
        for (int j = 0; j < S; ++j)
#pragma unroll
          for (int k = 0; k < S; ++k) wyz[j][k] = wy[j] * wz[k];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const double vs = (double)val[c * n + m] * scale;
#pragma unroll
          for (int i = 0; i < S; ++i) {
            const double u = odd_significand(wx[i] * vs);
#pragma unroll
            for (int j = 0; j < S; ++j)
#pragma unroll
              for (int k = 0; k < S; ++k) a[c][i][j][k] = __builtin_fma(u, wyz[j][k], a[c][i][j][k]);
          }
        }
      }
      // add the cell's integers into the tile: node (ax - OFF + i, ...) sits at LDS
      // (ax - OFF + i) - (x0 - OFF) after the bounded-axis clamp
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const int64_t nx = g.code[0] == 1 ? ax - OFF + i : wrap_index(ax - OFF + i, g.num[0], 0);
        const int ix = (int)(nx - (x0 - OFF));
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const int64_t ny = g.code[1] == 1 ? ay - OFF + j : wrap_index(ay - OFF + j, g.num[1], 0);
          const int iy = (int)(ny - (y0 - OFF));
#pragma unroll
          for (int k = 0; k < S; ++k) {
            const int64_t nz = g.code[2] == 1 ? az - OFF + k : wrap_index(az - OFF + k, g.num[2], 0);
            const int iz = (int)(nz - (z0 - OFF));
            const int at = (ix * LY + iy) * LZ + iz;
#pragma unroll
            for (int c = 0; c < C; ++c) {
              const long long q = biased_to_int(a[c][i][j][k]);
              if (q != 0) atomicAdd(&acc[c * L + at], (unsigned long long)q);
            }
          }
        }
      }
      done = stop;
    }
  }
  __syncthreads();

  // the tile into the field: plain read-modify-write, the colouring keeps tiles apart
  for (int i = threadIdx.x; i < L; i += BLOCK) {
    const int iz = i % LZ, iy = (i / LZ) % LY, ix = i / (LZ * LY);
    int64_t gx = x0 - OFF + ix, gy = y0 - OFF + iy, gz = z0 - OFF + iz;
    if (g.code[0] == 1) gx = wrap_index(gx, g.num[0], 1);
    if (g.code[1] == 1) gy = wrap_index(gy, g.num[1], 1);
    if (g.code[2] == 1) gz = wrap_index(gz, g.num[2], 1);
    if (gx < 0 || gx >= g.num[0] || gy < 0 || gy >= g.num[1] || gz < 0 || gz >= g.num[2]) continue;
    const int64_t at = ((gx * g.num[1] + gy) * g.num[2] + gz) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const long long q = (long long)acc[c * L + i];
      if (q != 0) field[at + c] += (double)q * inv;
    }
  }
}

// ---- tiled gather ---------------------------------------------------------------------

template <class T, int KIND, int P, int C, int TX, int TY, int TZ, int BLOCK>
__global__ void __launch_bounds__(BLOCK) gather_tiled(const T *__restrict__ pos, int64_t n,
                                                      const uint32_t *__restrict__ start, Grid g, int ntx, int nty,
                                                      const double *__restrict__ field, T *__restrict__ res) {
  using Sh = StencilShape<KIND, P>;
  constexpr int S = Sh::S, OFF = Sh::OFF;
  constexpr int LX = TX + S - 1, LY = TY + S - 1, LZ = TZ + S - 1, L = LX * LY * LZ;
  __shared__ double tile[C * L];

  const int tx = blockIdx.x % ntx, ty = (blockIdx.x / ntx) % nty, tz = blockIdx.x / (ntx * nty);
  const int64_t x0 = (int64_t)tx * TX, y0 = (int64_t)ty * TY, z0 = (int64_t)tz * TZ;
  for (int i = threadIdx.x; i < L; i += BLOCK) {
    const int iz = i % LZ, iy = (i / LZ) % LY, ix = i / (LZ * LY);
    // the footprint node, wrapped on a periodic axis and clamped (only reached by clamped
    // stencils) on a bounded one
    const int64_t gx = wrap_index(x0 - OFF + ix, g.num[0], g.code[0]);
    const int64_t gy = wrap_index(y0 - OFF + iy, g.num[1], g.code[1]);
    const int64_t gz = wrap_index(z0 - OFF + iz, g.num[2], g.code[2]);
    const int64_t at = ((gx * g.num[1] + gy) * g.num[2] + gz) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) tile[c * L + i] = field[at + c];
  }
  __syncthreads();

  for (int cl = threadIdx.x; cl < TX * TY * TZ; cl += BLOCK) {
    const int lz = cl % TZ, ly = (cl / TZ) % TY, lx = cl / (TZ * TY);
    const int64_t ax = x0 + lx, ay = y0 + ly, az = z0 + lz;
    if (!(ax < g.num[0] && ay < g.num[1] && az < g.num[2])) continue;
    const int64_t cell = (ax * g.num[1] + ay) * g.num[2] + az;
    const int64_t first = start[cell], last = start[cell + 1];
    if (first == last) continue;
    double f[C][S][S][S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int64_t nx = g.code[0] == 1 ? ax - OFF + i : wrap_index(ax - OFF + i, g.num[0], 0);
      const int ix = (int)(nx - (x0 - OFF));
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const int64_t ny = g.code[1] == 1 ? ay - OFF + j : wrap_index(ay - OFF + j, g.num[1], 0);
        const int iy = (int)(ny - (y0 - OFF));
#pragma unroll
        for (int k = 0; k < S; ++k) {
          const int64_t nz = g.code[2] == 1 ? az - OFF + k : wrap_index(az - OFF + k, g.num[2], 0);
          const int iz = (int)(nz - (z0 - OFF));
#pragma unroll
          for (int c = 0; c < C; ++c) f[c][i][j][k] = tile[c * L + (ix * LY + iy) * LZ + iz];
        }
      }
    }
    for (int64_t m = first; m < last; ++m) {
      double wx[4], wy[4], wz[4];
      anchor_axis<KIND, P>(((double)pos[m] - g.lower[0]) * g.inverse[0], g.num[0], g.code[0], wx);
      anchor_axis<KIND, P>(((double)pos[n + m] - g.lower[1]) * g.inverse[1], g.num[1], g.code[1], wy);
      anchor_axis<KIND, P>(((double)pos[2 * n + m] - g.lower[2]) * g.inverse[2], g.num[2], g.code[2], wz);
      double w[S][S][S];
#pragma unroll
      for (int i = 0; i < S; ++i)
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const double wab = wx[i] * wy[j];
#pragma unroll
          for (int k = 0; k < S; ++k) w[i][j][k] = wab * wz[k];
        }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        double total = 0.0;
#pragma unroll
        for (int i = 0; i < S; ++i)
#pragma unroll
          for (int j = 0; j < S; ++j)
#pragma unroll
            for (int k = 0; k < S; ++k) total += w[i][j][k] * f[c][i][j][k];
        res[c * n + m] = (T)total;
      }
    }
  }
}

// The dead bin (markers whose values were all zero at binning) is in no cell; a gather
// still interpolates at those positions, one marker per work-item, its range read from the
// offsets on the device.
template <class T, int KIND, int C>
__global__ void __launch_bounds__(256) gather_dead_kernel(const T *__restrict__ pos, int64_t n, int degree,
                                                          const uint32_t *__restrict__ start, int64_t cells, Grid g,
                                                          const double *__restrict__ field, T *__restrict__ res) {
  const int64_t first = start[cells], last = start[cells + 1];
  for (int64_t m = first + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < last;
       m += (int64_t)gridDim.x * blockDim.x) {
    int64_t node[3][4];
    double w[3][4];
    int size[3];
    for (int d = 0; d < 3; ++d)
      size[d] = axis_stencil<KIND>(((double)pos[d * n + m] - g.lower[d]) * g.inverse[d], g.num[d], g.code[d], degree,
                                   node[d], w[d]);
    for (int c = 0; c < C; ++c) {
      double total = 0.0;
      for (int a = 0; a < size[0]; ++a)
        for (int b = 0; b < size[1]; ++b) {
          const double wab = w[0][a] * w[1][b];
          for (int e = 0; e < size[2]; ++e)
            total += wab * w[2][e] * field[((node[0][a] * g.num[1] + node[1][b]) * g.num[2] + node[2][e]) * C + c];
        }
      res[c * n + m] = (T)total;
    }
  }
}

// ---- host side ------------------------------------------------------------------------

struct TileShape {
  int tx, ty, tz, block;
};

// Tile extents per (kind, degree, ncomp); LDS holds C x (T + S - 1)^3 int64 / float64.
static TileShape tile_shape(int kind, int degree, int ncomp) {
  (void)kind;
  (void)degree;
  (void)ncomp;
  return TileShape{8, 8, 16, 256};
}

static int colors_for(int nt, int code) {
  if (nt == 1) return 1;
  if (code == 1 && (nt % 2) == 1) return 3;
  return 2;
}

// Which (kind, degree, ncomp, dims, codes) the tiled kernels take; others use the direct path.
static bool tiled_supported(int kind, int degree, int ndim, int64_t ncomp, const Grid &g, const TileShape &ts) {
  if (ndim != 3) return false;
  const int p = kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree);
  if (!((ncomp == 1 || ncomp == 3) && (p == 1 || p == 2 || (p == 3 && ncomp == 1)))) return false;
  if (kind == SPLINE && degree != 3) return false;  // linear / quadratic have their own kinds
  const int t[3] = {ts.tx, ts.ty, ts.tz};
  for (int d = 0; d < 3; ++d) {
    if (g.code[d] == 2) return false;
    // a periodic axis needs at least two tiles so that a footprint never meets itself
    if (g.code[d] == 1 && (g.num[d] + t[d] - 1) / t[d] < 2) return false;
    if (g.num[d] < t[d]) return false;
    // and whole tiles, so that same-colour footprints stay apart across the seam
    if (g.code[d] == 1 && g.num[d] % t[d] != 0) return false;
  }
  return true;
}

}  // namespace mpf

using namespace mpf;

// The bins a caller keeps: sorted copies of the markers, the permutation and the cell
// offsets.  All device memory owned by the caller; sizes from mpf_bin_sizes.
extern "C" int mpf_bin_sizes(int ndim, const int64_t *num_nodes, int64_t *cells) {
  if (ndim < 1 || ndim > 3) {
    set_error("ndim must be 1, 2 or 3");
    return MPF_ERR_ARGUMENT;
  }
  int64_t c = 1;
  for (int d = 0; d < ndim; ++d) c *= num_nodes[d];
  *cells = c;  // the offsets array holds cells + 2 entries (the dead bin last)
  return MPF_OK;
}

template <class T, int KIND, int P, int NDIM>
static void launch_count(const void *pos, int64_t ps, const void *val, int64_t vs, int64_t n, int ncomp, const Grid &g,
                         int64_t dead_key, uint32_t *key, uint32_t *rank, uint32_t *count, unsigned long long *maxbits,
                         hipStream_t stream) {
  hipLaunchKernelGGL((bin_count_kernel<T, KIND, P, NDIM>), dim3(blocks_for(n, 256, 1 << 16)), dim3(256), 0, stream,
                     (const T *)pos, ps, (const T *)val, vs, n, ncomp, g, dead_key, key, rank, count, maxbits);
}

template <class T, int KIND, int P>
static void dispatch_count_nd(int ndim, const void *pos, int64_t ps, const void *val, int64_t vs, int64_t n, int ncomp,
                              const Grid &g, int64_t dead_key, uint32_t *key, uint32_t *rank, uint32_t *count,
                              unsigned long long *maxbits, hipStream_t stream) {
  if (ndim == 1) launch_count<T, KIND, P, 1>(pos, ps, val, vs, n, ncomp, g, dead_key, key, rank, count, maxbits, stream);
  if (ndim == 2) launch_count<T, KIND, P, 2>(pos, ps, val, vs, n, ncomp, g, dead_key, key, rank, count, maxbits, stream);
  if (ndim == 3) launch_count<T, KIND, P, 3>(pos, ps, val, vs, n, ncomp, g, dead_key, key, rank, count, maxbits, stream);
}

template <class T>
static int dispatch_count(int kind, int degree, int ndim, const void *pos, int64_t ps, const void *val, int64_t vs,
                          int64_t n, int ncomp, const Grid &g, int64_t dead_key, uint32_t *key, uint32_t *rank,
                          uint32_t *count, unsigned long long *maxbits, hipStream_t stream) {
#define CNT(K, P) dispatch_count_nd<T, K, P>(ndim, pos, ps, val, vs, n, ncomp, g, dead_key, key, rank, count, maxbits, stream)
  if (kind == LINEAR) CNT(LINEAR, 1);
  else if (kind == QUADRATIC) CNT(QUADRATIC, 2);
  else if (degree == 0) CNT(SPLINE, 0);
  else if (degree == 1) CNT(SPLINE, 1);
  else if (degree == 2) CNT(SPLINE, 2);
  else CNT(SPLINE, 3);
#undef CNT
  return MPF_OK;
}

/* Bin markers by stencil cell (row-major anchor order).  Writes sorted positions
 * (ndim, n) and values (ncomp, n) with contiguous rows, the permutation order[sorted] =
 * original index (int64), the cell offsets start[cells + 2] (uint32; start[cells] is the
 * first dead marker, start[cells + 1] = n) and, in scalars[0..1], the fixed-point scale
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
  if (n < 0 || n > 0x7FFFFFFFll || ncomp < 0 || ncomp > 64 || (ncomp > 0 && !values)) {
    set_error("bin_markers: need 0 <= n < 2^31 and 0 <= ncomp <= 64 (0 only without values)");
    return MPF_ERR_ARGUMENT;
  }
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size, num_nodes,
                     code);
  if (rc) return rc;
  for (int d = 0; d < ndim; ++d)
    if (code[d] == 2) {
      set_error("bin_markers: clamped axes are not binned (the direct kernels take them)");
      return MPF_ERR_UNSUPPORTED;
    }
  int64_t cells = 1;
  for (int d = 0; d < ndim; ++d) cells *= num_nodes[d];
  if (cells + 2 > 0x7FFFFFFFll) {
    set_error("bin_markers: too many cells for 32-bit keys");
    return MPF_ERR_ARGUMENT;
  }
  hipStream_t stream = (hipStream_t)stream_;
  if (!ctx) ctx = default_context();
  if (!ctx) return MPF_ERR_HIP;
  // scratch: key + rank per marker, the counts, the scan's temporary storage
  rc = ensure(ctx->sort_keys, (size_t)(n > 0 ? n : 1) * 8, false, stream);
  if (rc) return rc;
  uint32_t *key = (uint32_t *)ctx->sort_keys.ptr, *rank = key + (n > 0 ? n : 1);
  rc = ensure(ctx->offsets, (size_t)(cells + 2) * 4, false, stream);
  if (rc) return rc;
  uint32_t *count = (uint32_t *)ctx->offsets.ptr;
  rc = zero_async(count, (size_t)(cells + 2) * 4, stream);
  if (rc) return rc;
  rc = ensure(ctx->scalars, 64, true, stream);
  if (rc) return rc;
  unsigned long long *maxbits = (unsigned long long *)ctx->scalars.ptr;
  rc = zero_async(maxbits, 8, stream);
  if (rc) return rc;
  if (n > 0) {
    if (dtype == F32)
      dispatch_count<float>(kind, degree, ndim, positions, pos_stride, values, val_stride, n, (int)ncomp, g, cells,
                            key, rank, count, maxbits, stream);
    else
      dispatch_count<double>(kind, degree, ndim, positions, pos_stride, values, val_stride, n, (int)ncomp, g, cells,
                             key, rank, count, maxbits, stream);
    MPF_LAUNCH_CHECK();
  }
  size_t temp = 0;
  MPF_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, count, start, (int)(cells + 2), stream));
  rc = ensure(ctx->sort_tmp, temp, false, stream);
  if (rc) return rc;
  MPF_CHECK(hipcub::DeviceScan::ExclusiveSum(ctx->sort_tmp.ptr, temp, count, start, (int)(cells + 2), stream));
  if (n > 0) {
    if (dtype == F32)
      hipLaunchKernelGGL(bin_place_kernel<float>, dim3(blocks_for(n, 256, 1 << 16)), dim3(256), 0, stream,
                         (const float *)positions, pos_stride, ndim, (const float *)values, val_stride, (int)ncomp, n,
                         key, rank, start, (float *)sorted_positions, (float *)sorted_values, order);
    else
      hipLaunchKernelGGL(bin_place_kernel<double>, dim3(blocks_for(n, 256, 1 << 16)), dim3(256), 0, stream,
                         (const double *)positions, pos_stride, ndim, (const double *)values, val_stride, (int)ncomp,
                         n, key, rank, start, (double *)sorted_positions, (double *)sorted_values, order);
    MPF_LAUNCH_CHECK();
  }
  // the scale, from the maximum the count pass found
  rc = prepare_scale(ctx, nullptr, 0, dtype, 1, n, 1.0, scale_exp, stream);
  if (rc) return rc;
  MPF_CHECK(hipMemcpyAsync(scalars, scale_slot(ctx), 24, hipMemcpyDeviceToDevice, stream));
  return MPF_OK;
}

template <class T, int KIND, int P, int C>
static int run_scatter_tiled(const void *pos, const void *val, int64_t n, const uint32_t *start, const Grid &g,
                             double *field, const double *scalars, hipStream_t stream) {
  constexpr int TX = 8, TY = 8, TZ = 16, BLOCK = 256;
  Tiling tl;
  tl.tx = TX;
  tl.ty = TY;
  tl.tz = TZ;
  tl.ntx = (int)((g.num[0] + TX - 1) / TX);
  tl.nty = (int)((g.num[1] + TY - 1) / TY);
  tl.ntz = (int)((g.num[2] + TZ - 1) / TZ);
  tl.cx = colors_for(tl.ntx, g.code[0]);
  tl.cy = colors_for(tl.nty, g.code[1]);
  tl.cz = colors_for(tl.ntz, g.code[2]);
  auto count_of = [](int nt, int colors, int color) {
    if (colors == 1) return nt;
    if (colors == 3) return color == 2 ? 1 : (nt - 1 + 1 - color) / 2;
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
        hipLaunchKernelGGL((scatter_tiled<T, KIND, P, C, TX, TY, TZ, BLOCK>), dim3(blocks), dim3(BLOCK), 0, stream,
                           (const T *)pos, (const T *)val, n, start, g, tl, field, scalars);
        MPF_LAUNCH_CHECK();
      }
  return MPF_OK;
}

template <class T, int KIND, int P, int C>
static int run_gather_tiled(const void *pos, int64_t n, const uint32_t *start, const Grid &g, const double *field,
                            void *res, hipStream_t stream) {
  constexpr int TX = 8, TY = 8, TZ = 16, BLOCK = 256;
  const int ntx = (int)((g.num[0] + TX - 1) / TX), nty = (int)((g.num[1] + TY - 1) / TY),
            ntz = (int)((g.num[2] + TZ - 1) / TZ);
  {
    ProfileScope prof("gather_tiled", stream);
    hipLaunchKernelGGL((gather_tiled<T, KIND, P, C, TX, TY, TZ, BLOCK>), dim3(ntx * nty * ntz), dim3(BLOCK), 0,
                       stream, (const T *)pos, n, start, g, ntx, nty, field, (T *)res);
  }
  MPF_LAUNCH_CHECK();
  const int64_t cells = g.num[0] * g.num[1] * g.num[2];
  hipLaunchKernelGGL((gather_dead_kernel<T, KIND, C>), dim3(64), dim3(256), 0, stream, (const T *)pos, n, P, start,
                     cells, g, field, (T *)res);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

#define TILED_CASES(X, T)                                \
  if (kind == LINEAR && ncomp == 1) return X(T, LINEAR, 1, 1);       \
  if (kind == LINEAR && ncomp == 3) return X(T, LINEAR, 1, 3);       \
  if (kind == QUADRATIC && ncomp == 1) return X(T, QUADRATIC, 2, 1); \
  if (kind == QUADRATIC && ncomp == 3) return X(T, QUADRATIC, 2, 3); \
  if (kind == SPLINE && degree == 3 && ncomp == 1) return X(T, SPLINE, 3, 1);

/* Deposit binned markers into field with the tiled kernels (int64 fixed point through
 * float64 exact accumulators; bit-reproducible).  sorted_* and start come from
 * mpf_bin_markers of the same kind/degree/grid; scalars holds [scale, 1/scale].
 * Returns MPF_ERR_UNSUPPORTED for shapes the tiled kernels do not cover. */
extern "C" int mpf_scatter_binned(int kind, int degree, int ndim, const void *sorted_positions,
                                  const void *sorted_values, int dtype, int64_t n, int64_t ncomp,
                                  const uint32_t *start, const double *scalars, double *field, const double *lower,
                                  const double *cell_size, const int64_t *num_nodes, const int64_t *code,
                                  mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size,
                     num_nodes, code);
  if (rc) return rc;
  const TileShape ts = tile_shape(kind, degree, ncomp);
  if (!tiled_supported(kind, degree, ndim, ncomp, g, ts)) {
    set_error("no tiled kernel for kind %d degree %d, %dD, %lld components on this grid", kind, degree, ndim,
              (long long)ncomp);
    return MPF_ERR_UNSUPPORTED;
  }
  hipStream_t stream = (hipStream_t)stream_;
#define SC(T, K, P, C) run_scatter_tiled<T, K, P, C>(sorted_positions, sorted_values, n, start, g, field, scalars, stream)
  if (dtype == F32) {
    TILED_CASES(SC, float)
  } else {
    TILED_CASES(SC, double)
  }
#undef SC
  set_error("no tiled kernel for this case");
  return MPF_ERR_UNSUPPORTED;
}

/* Gather at binned markers with the tiled kernels: result (ncomp, n) in the sorted order. */
extern "C" int mpf_gather_binned(int kind, int degree, int ndim, const void *sorted_positions, int dtype, int64_t n,
                                 const uint32_t *start, const double *field, int64_t ncomp, void *result,
                                 const double *lower, const double *cell_size, const int64_t *num_nodes,
                                 const int64_t *code, mpf_context *ctx, mpf_stream stream_) {
  (void)ctx;
  Grid g;
  int rc = make_grid(g, ndim, kind, kind == LINEAR ? 1 : (kind == QUADRATIC ? 2 : degree), lower, cell_size,
                     num_nodes, code);
  if (rc) return rc;
  const TileShape ts = tile_shape(kind, degree, ncomp);
  if (!tiled_supported(kind, degree, ndim, ncomp, g, ts)) {
    set_error("no tiled kernel for kind %d degree %d, %dD, %lld components on this grid", kind, degree, ndim,
              (long long)ncomp);
    return MPF_ERR_UNSUPPORTED;
  }
  hipStream_t stream = (hipStream_t)stream_;
#define GA(T, K, P, C) run_gather_tiled<T, K, P, C>(sorted_positions, n, start, g, field, result, stream)
  if (dtype == F32) {
    TILED_CASES(GA, float)
  } else {
    TILED_CASES(GA, double)
  }
#undef GA
  set_error("no tiled kernel for this case");
  return MPF_ERR_UNSUPPORTED;
}
