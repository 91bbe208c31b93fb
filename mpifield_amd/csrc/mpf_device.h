// mpf_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Per-axis stencils with the arithmetic of the reference's Pyccel kernels (cell selection
// in FP64: multiply by the inverse spacing, floor modulo only outside [0, n), int()
// truncation), so that every marker lands in the same cell as on the CPU:
//   linear     linear_cell_pyccel.py:22-57
//   quadratic  scatter_quadratic_3d_pyccel.py:28-54
//   spline     spline_cell_pyccel.py:165-218 (degree 0..3), clamped_cell :106-162
// All files are compiled with -ffp-contract=off: a multiply-add is fused only where a
// kernel writes fma() itself.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpf {

enum Kind : int { LINEAR = 0, QUADRATIC = 1, SPLINE = 2 };
enum Mode : int { MODE_ATOMIC = 0, MODE_FIXED = 1 };
enum Dtype : int { F32 = 0, F64 = 1 };

// The grid of one block as the kernels see it (BlockGrid.geometry, _types.py:27-44).
struct Grid {
  int ndim;
  int code[3];         // 0 bounded (fold), 1 periodic, 2 clamped spline knots
  int64_t num[3];      // nodes per axis
  int64_t extent[3];   // array extent per axis: num, or num + degree - 1 on a clamped axis
  int64_t stride[3];   // node stride per axis in the C-ordered array (components excluded)
  double lower[3];
  double cell[3];
  double inverse[3];   // 1.0 / cell, as the reference computes it once per call
};

// k modulo num for a stencil node of a wrapped position (at most 3 nodes outside
// [0, num)): on an axis of 8 nodes or more one conditional add or subtract, instead of the
// 64-bit division a remainder costs; the remainder itself on tinier axes.
__device__ __forceinline__ int64_t wrap_near(int64_t k, int64_t num) {
  if (num < 8) return ((k % num) + num) % num;
  k = k < 0 ? k + num : k;
  return k >= num ? k - num : k;
}

__device__ __forceinline__ double periodic_position(double rel, int64_t num) {
  const double n = (double)num;
  return (rel >= 0.0 && rel < n) ? rel : rel - n * floor(rel / n);
}

// Linear: left node, right node and the fraction of the right one.
__device__ __forceinline__ int linear_axis(double rel, int64_t num, int code, int64_t *node, double *w) {
  double r = rel;
  int64_t l;
  if (code == 1) {
    r = periodic_position(r, num);
    int64_t i = (int64_t)r;
    l = i >= num ? num - 1 : i;
  } else {
    int64_t i = r >= 0.0 ? (int64_t)r : 0;
    l = i > num - 2 ? num - 2 : i;
  }
  node[0] = l;
  node[1] = (code == 1 && l + 1 == num) ? 0 : l + 1;
  double d = r - (double)l;
  double f = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
  w[0] = 1.0 - f;
  w[1] = f;
  return 2;
}

__device__ __forceinline__ int quadratic_axis(double rel, int64_t num, int code, int64_t *node, double *w) {
  double r = rel;
  int64_t c;
  if (code == 1) {
    if (!(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
    c = (int64_t)(r + 0.5);
    node[0] = wrap_near(c - 1, num);
    node[1] = wrap_near(c, num);
    node[2] = wrap_near(c + 1, num);
  } else {
    if (r < 0.0) r = 0.0;
    if (r > (double)(num - 1)) r = (double)num - 1.0;
    c = (int64_t)(r + 0.5);
    node[0] = c > 0 ? c - 1 : 0;
    node[1] = c;
    node[2] = c < num - 1 ? c + 1 : num - 1;
  }
  double o = r - (double)c;
  w[0] = 0.5 * (0.5 - o) * (0.5 - o);
  w[1] = 0.75 - o * o;
  w[2] = 0.5 * (0.5 + o) * (0.5 + o);
  return 3;
}

// Cox-de Boor on the clamped knots t_k = clip(k - p, 0, num - 1); staggered 1 gives the
// rescaled degree-(p-1) functions of the derivative identity.
__device__ inline void clamped_cell(double rel, int64_t num, int degree, int staggered, int64_t *node, double *w) {
  double left[4] = {0, 0, 0, 0}, right[4] = {0, 0, 0, 0};
  double r = rel;
  if (r < 0.0) r = 0.0;
  if (r > (double)(num - 1)) r = (double)num - 1.0;
  int64_t cell = (int64_t)r;
  if (cell > num - 2) cell = num - 2;
  const int64_t span = cell + degree;
  const int order = degree - staggered;
  w[0] = 1.0;
  for (int k = 1; k <= order; ++k) {
    int64_t low = span + 1 - k - degree, high = span + k - degree;
    if (low < 0) low = 0;
    if (high > num - 1) high = num - 1;
    left[k] = r - (double)low;
    right[k] = (double)high - r;
    double saved = 0.0;
    for (int m = 0; m < k; ++m) {
      double temp = w[m] / (right[m + 1] + left[k - m]);
      w[m] = saved + right[m + 1] * temp;
      saved = left[k - m] * temp;
    }
    w[k] = saved;
  }
  for (int m = 0; m <= order; ++m) {
    node[m] = cell + m;
    if (staggered == 1) {
      int64_t i = cell + 1 + m, low = i - degree, high = i;
      if (low < 0) low = 0;
      if (high > num - 1) high = num - 1;
      w[m] = w[m] * (double)degree / (double)(high - low);
    }
  }
}

__device__ inline int spline_axis(double rel, int64_t num, int code, int degree, int64_t *node, double *w) {
  if (code == 2) {
    clamped_cell(rel, num, degree, 0, node, w);
    return degree + 1;
  }
  double r = rel;
  if (code == 1) {
    if (!(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
  } else {
    if (r < 0.0) r = 0.0;
    if (r > (double)(num - 1)) r = (double)num - 1.0;
  }
  const int64_t middle = (degree % 2 == 0) ? (int64_t)(r + 0.5) : (int64_t)r;
  const double o = r - (double)middle;
  if (degree == 0) {
    w[0] = 1.0;
  } else if (degree == 1) {
    w[0] = 1.0 - o;
    w[1] = o;
  } else if (degree == 2) {
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
  const int64_t first = middle - degree / 2;
  for (int a = 0; a <= degree; ++a) {
    int64_t k = first + a;
    if (code == 1) {
      k = wrap_near(k, num);
    } else {
      if (k < 0) k = 0;
      if (k > num - 1) k = num - 1;
    }
    node[a] = k;
  }
  return degree + 1;
}

__device__ __forceinline__ int64_t floor_int(double x) {
  int64_t whole = (int64_t)x;
  if (x < (double)whole) whole -= 1;
  return whole;
}

// spline_edges_pyccel.py:86-138: the `degree` edges along an axis around rel, weights
// B_{p-1}(rel - k - 1/2); on a bounded axis an edge past a wall does not exist.
__device__ __forceinline__ void along_stencil(double rel, int64_t num, int code, int degree, int64_t *edge, double *w,
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
      edge[a] = wrap_near(k, num);
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

template <int KIND>
__device__ __forceinline__ int axis_stencil(double rel, int64_t num, int code, int degree, int64_t *node, double *w) {
  if (KIND == LINEAR) return linear_axis(rel, num, code, node, w);
  if (KIND == QUADRATIC) return quadratic_axis(rel, num, code, node, w);
  return spline_axis(rel, num, code, degree, node, w);
}

// d/dx of the stencil weights along one axis, in cells (times the inverse spacing by the
// caller).  Zero where the position is clamped to a bounded wall or for degree 0.
template <int KIND>
__device__ inline int axis_derivative(double rel, int64_t num, int code, int degree, int64_t *node, double *w,
                                      double *dw) {
  const int size = axis_stencil<KIND>(rel, num, code, degree, node, w);
  const bool clamped = code == 0 && (rel < 0.0 || rel > (double)(num - 1));
  const int p = KIND == LINEAR ? 1 : (KIND == QUADRATIC ? 2 : degree);
  if (clamped || p == 0) {
    for (int a = 0; a < size; ++a) dw[a] = 0.0;
    return size;
  }
  if (p == 1) {
    // linear and spline p=1: w = (1 - f, f), f = rel - left within the cell
    dw[0] = -1.0;
    dw[1] = 1.0;
    if (KIND == LINEAR && code == 0 && rel > (double)(num - 1)) dw[0] = dw[1] = 0.0;
    return size;
  }
  // the offset from the middle node, recomputed as the stencil did
  double r = rel;
  if (code == 1 && !(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
  const double o = p == 2 ? r - (double)(int64_t)(r + 0.5) : r - (double)(int64_t)r;
  if (p == 2) {
    dw[0] = -(0.5 - o);
    dw[1] = -2.0 * o;
    dw[2] = 0.5 + o;
  } else {
    dw[0] = -0.5 * (1.0 - o) * (1.0 - o);
    dw[1] = 1.5 * o * o - 2.0 * o;
    dw[2] = -1.5 * o * o + o + 0.5;
    dw[3] = 0.5 * o * o;
  }
  return size;
}

template <class T>
__device__ __forceinline__ double load_as_double(const T *p) {
  return (double)(*p);
}

// Fixed point: the nearest integer of x, in two's complement (|x| < 2^62 by the choice of scale).
__device__ __forceinline__ unsigned long long to_fixed(double x) {
  return (unsigned long long)__double2ll_rn(x);
}

// The same integer for |x| < 2^51 in two float64/integer operations: x + 1.5 * 2^52 lies in
// [2^52, 2^53), where the unit in the last place is 1, so the add rounds x to the nearest
// integer (ties to even, as __double2ll_rn), and the sum's bits less the bias's bits are
// that integer in two's complement.
__device__ __forceinline__ unsigned long long to_fixed_small(double x) {
  const double bias = 6755399441055744.0;  // 1.5 * 2^52
  return (unsigned long long)__double_as_longlong(x + bias) - (unsigned long long)__double_as_longlong(bias);
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

}  // namespace mpf
