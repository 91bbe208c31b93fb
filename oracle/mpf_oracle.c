/*
 * mpf_oracle.c -- the CPU oracle of the particle-grid hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C, FP64 restatement of the
 * reference's Pyccel kernels (max-models/mpifield, src/mpifield/kernels/<name>/<name>_pyccel.py).
 * It is the checker the GPU library is compared against, and the timed "host path"
 * of the benchmark.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never calls into it.
 *
 * Each function cites the reference lines it follows.  The arithmetic is kept in the
 * reference's order (multiply by the inverse spacing, the floor modulo only outside
 * [0, n), int() truncation, weight products left to right); it is compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Thread policy (the reference's OpenMP form):
 *   node deposits  serial if threads == 1, N < 20000, N < grid values, or
 *                  threads * grid values > 16,777,216; otherwise each thread
 *                  deposits a contiguous chunk into its own zeroed copy of the grid
 *                  and the copies are summed in thread order
 *                  (scatter_quadratic_3d_pyccel.py:150-218).
 *   linear 2D/3D   the reference uses x-banded buffers (scatter_linear_2d_pyccel.py:
 *                  192-261); here full-grid copies, which differ from the bands only
 *                  in the order of the per-node sum.
 *   gathers        independent per marker (threads only split the loop).
 *   Whitney paths  per-thread copies of the edges when N >= 20000 and N >= nodes,
 *                  added thread by thread (deposit_paths_2d_pyccel.py:155-183).
 *   spline paths   serial (deposit_paths_spline_2d_pyccel.py:7).
 *   forms          deposits as the node deposits (scatter_forms_spline_3d_pyccel.py:
 *                  153-210); gathers split into contiguous chunks.
 *
 * Parity pins (tests/test_oracle.py): known answers of the reference tests, closed
 * forms, scipy's BSpline for clamped knots, the adjoint identity, charge totals,
 * continuity G^T j = rho(end) - rho(start), halo fold/fill = bare periodic grid, the
 * work identity j.e = sum q int E.dl for the edge gathers, the commuting property of the
 * forms (the (k+1)-form field of the discrete exterior derivative is the derivative of the
 * k-form field), and the reference's own C kernels (scatter_linear_1d_c.c,
 * scatter_rows_3d_c.c) built into oracle/_ref.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MIN_PARALLEL_SCATTER 20000
#define MIN_PARALLEL_GATHER 4096
#define MAX_BUFFER_VALUES 16777216

enum { KIND_LINEAR = 0, KIND_QUADRATIC = 1, KIND_SPLINE = 2 };

/* ------------------------------------------------------------------------- */
/* Per-axis stencils                                                          */
/* ------------------------------------------------------------------------- */

/* linear_cell_pyccel.py:22-28 periodic_position */
static inline double periodic_position(double rel, int64_t num) {
    return (rel >= 0.0 && rel < (double)num) ? rel : rel - (double)num * floor(rel / (double)num);
}

/* linear_cell_pyccel.py:31-45 left_node */
static inline int64_t left_node(double rel, int64_t num, int64_t periodic) {
    if (periodic == 1) {
        int64_t i = (int64_t)rel;
        return i >= num ? num - 1 : i;
    }
    int64_t i = rel >= 0.0 ? (int64_t)rel : 0;
    return i > num - 2 ? num - 2 : i;
}

/* linear_cell_pyccel.py:48-51 right_node */
static inline int64_t right_node(int64_t left, int64_t num, int64_t periodic) {
    return (periodic == 1 && left + 1 == num) ? 0 : left + 1;
}

/* linear_cell_pyccel.py:54-57 cell_fraction */
static inline double cell_fraction(double rel, int64_t left) {
    double d = rel - (double)left;
    return d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
}

/* scatter_quadratic_3d_pyccel.py:28-54 _quadratic_cell */
static void quadratic_cell(double rel, int64_t num, int64_t periodic, int64_t node[3], double weight[3]) {
    double r = rel;
    int64_t centre;
    if (periodic == 1) {
        if (!(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
        centre = (int64_t)(r + 0.5);
        node[0] = (centre + num - 1) % num;
        node[1] = centre % num;
        node[2] = (centre + 1) % num;
    } else {
        if (r < 0.0) r = 0.0;
        if (r > (double)(num - 1)) r = (double)num - 1.0;
        centre = (int64_t)(r + 0.5);
        node[0] = centre > 0 ? centre - 1 : 0;
        node[1] = centre;
        node[2] = centre < num - 1 ? centre + 1 : num - 1;
    }
    double offset = r - (double)centre;
    weight[0] = 0.5 * (0.5 - offset) * (0.5 - offset);
    weight[1] = 0.75 - offset * offset;
    weight[2] = 0.5 * (0.5 + offset) * (0.5 + offset);
}

/* spline_cell_pyccel.py:106-162 clamped_cell (Cox-de Boor on clamped knots) */
static void clamped_cell(double rel, int64_t num, int degree, int staggered, int64_t *node, double *weight) {
    double left[4] = {0, 0, 0, 0}, right[4] = {0, 0, 0, 0};
    double r = rel;
    if (r < 0.0) r = 0.0;
    if (r > (double)(num - 1)) r = (double)num - 1.0;
    int64_t cell = (int64_t)r;
    if (cell > num - 2) cell = num - 2;
    int64_t span = cell + degree;
    int order = degree - staggered;
    weight[0] = 1.0;
    for (int k = 1; k <= order; ++k) {
        int64_t low = span + 1 - k - degree, high = span + k - degree;
        if (low < 0) low = 0;
        if (high > num - 1) high = num - 1;
        left[k] = r - (double)low;
        right[k] = (double)high - r;
        double saved = 0.0;
        for (int m = 0; m < k; ++m) {
            double temp = weight[m] / (right[m + 1] + left[k - m]);
            weight[m] = saved + right[m + 1] * temp;
            saved = left[k - m] * temp;
        }
        weight[k] = saved;
    }
    for (int m = 0; m <= order; ++m) {
        node[m] = cell + m;
        if (staggered == 1) {
            int64_t i = cell + 1 + m, low = i - degree, high = i;
            if (low < 0) low = 0;
            if (high > num - 1) high = num - 1;
            weight[m] = weight[m] * (double)degree / (double)(high - low);
        }
    }
}

/* spline_cell_pyccel.py:165-218 spline_cell */
static void spline_cell(double rel, int64_t num, int64_t periodic, int degree, int64_t *node, double *weight) {
    if (periodic == 2) {
        clamped_cell(rel, num, degree, 0, node, weight);
        return;
    }
    double r = rel;
    if (periodic == 1) {
        if (!(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
    } else {
        if (r < 0.0) r = 0.0;
        if (r > (double)(num - 1)) r = (double)num - 1.0;
    }
    int64_t middle = (degree % 2 == 0) ? (int64_t)(r + 0.5) : (int64_t)r;
    double offset = r - (double)middle;
    if (degree == 0) {
        weight[0] = 1.0;
    } else if (degree == 1) {
        weight[0] = 1.0 - offset;
        weight[1] = offset;
    } else if (degree == 2) {
        weight[0] = 0.5 * (0.5 - offset) * (0.5 - offset);
        weight[1] = 0.75 - offset * offset;
        weight[2] = 0.5 * (0.5 + offset) * (0.5 + offset);
    } else {
        double complement = 1.0 - offset, square = offset * offset, cube = square * offset;
        weight[0] = complement * complement * complement / 6.0;
        weight[1] = (3.0 * cube - 6.0 * square + 4.0) / 6.0;
        weight[2] = (-3.0 * cube + 3.0 * square + 3.0 * offset + 1.0) / 6.0;
        weight[3] = cube / 6.0;
    }
    int64_t first = middle - degree / 2;
    for (int a = 0; a <= degree; ++a) {
        int64_t k = first + a;
        if (periodic == 1) {
            k = (k + num) % num;
        } else {
            if (k < 0) k = 0;
            if (k > num - 1) k = num - 1;
        }
        node[a] = k;
    }
}

/* One axis of a node weighting: the stencil of rel on an axis of num nodes.
 * Linear: left/right node and (1 - f, f) (gather_linear_2d_pyccel.py:49-54).
 * Returns the stencil size. */
static int axis_stencil(int kind, int degree, double rel, int64_t num, int64_t code, int64_t *node, double *weight) {
    if (kind == KIND_LINEAR) {
        double r = rel;
        if (code == 1) r = periodic_position(r, num);
        int64_t l = left_node(r, num, code);
        node[0] = l;
        node[1] = right_node(l, num, code);
        double f = cell_fraction(r, l);
        weight[0] = 1.0 - f;
        weight[1] = f;
        return 2;
    }
    if (kind == KIND_QUADRATIC) {
        quadratic_cell(rel, num, code, node, weight);
        return 3;
    }
    spline_cell(rel, num, code, degree, node, weight);
    return degree + 1;
}

/* ------------------------------------------------------------------------- */
/* Node deposits and gathers                                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    int ndim;
    double lower[3], inverse[3], cell[3];
    int64_t num[3], code[3];
    int64_t extent[3]; /* array extent per axis: num (+ degree - 1 on a clamped axis) */
} grid_t;

static void make_grid(grid_t *g, int ndim, int kind, int degree, const double *lower, const double *cell,
                      const int64_t *num, const int64_t *code) {
    g->ndim = ndim;
    for (int d = 0; d < 3; ++d) {
        g->lower[d] = d < ndim ? lower[d] : 0.0;
        g->cell[d] = d < ndim ? cell[d] : 1.0;
        g->inverse[d] = 1.0 / g->cell[d];
        g->num[d] = d < ndim ? num[d] : 1;
        g->code[d] = d < ndim ? code[d] : 0;
        g->extent[d] = g->num[d];
        if (d < ndim && kind == KIND_SPLINE && code[d] == 2) g->extent[d] = num[d] + degree - 1;
    }
}

/* The deposit of markers [first, last) into buf (a grid of the field's shape).
 * 1D linear: scatter_linear_1d_pyccel.py:27-54 (divides by the cell size).
 * 2D linear: scatter_linear_2d_pyccel.py:35-111.  3D linear: scatter_linear_3d_pyccel.py:37-171.
 * Quadratic/spline: scatter_quadratic_3d_pyccel.py:57-127, scatter_spline_3d_pyccel.py:29-104. */
static void scatter_chunk(int kind, int degree, const grid_t *g, const double *pos, int64_t ps, const double *val,
                          int64_t vs, int64_t ncomp, double *buf, int64_t first, int64_t last) {
    const int ndim = g->ndim;
    const int64_t s1 = g->extent[1] * g->extent[2], s2 = g->extent[2];
    for (int64_t m = first; m < last; ++m) {
        int live = 0;
        for (int64_t c = 0; c < ncomp; ++c)
            if (val[c * vs + m] != 0.0) live = 1;
        if (!live) continue;
        if (kind == KIND_LINEAR && ndim == 1) {
            /* one component at a time, each its own kernel call in the reference */
            double rel = (pos[m] - g->lower[0]) / g->cell[0];
            if (g->code[0] == 1) rel = periodic_position(rel, g->num[0]);
            int64_t l = left_node(rel, g->num[0], g->code[0]);
            int64_t r = right_node(l, g->num[0], g->code[0]);
            double f = cell_fraction(rel, l);
            for (int64_t c = 0; c < ncomp; ++c) {
                double v = val[c * vs + m];
                if (v == 0.0) continue; /* per-component kernel skips its zero values */
                buf[l * ncomp + c] += (1.0 - f) * v;
                buf[r * ncomp + c] += f * v;
            }
            continue;
        }
        if (kind == KIND_LINEAR && ndim == 2) {
            double rx = (pos[m] - g->lower[0]) * g->inverse[0];
            double ry = (pos[ps + m] - g->lower[1]) * g->inverse[1];
            if (g->code[0] == 1) rx = periodic_position(rx, g->num[0]);
            int64_t il = left_node(rx, g->num[0], g->code[0]);
            int64_t ir = right_node(il, g->num[0], g->code[0]);
            if (g->code[1] == 1) ry = periodic_position(ry, g->num[1]);
            int64_t jl = left_node(ry, g->num[1], g->code[1]);
            int64_t jr = right_node(jl, g->num[1], g->code[1]);
            double fx = cell_fraction(rx, il), fy = cell_fraction(ry, jl);
            double wll = (1.0 - fx) * (1.0 - fy), wlr = fx * (1.0 - fy), wul = (1.0 - fx) * fy, wur = fx * fy;
            for (int64_t c = 0; c < ncomp; ++c) {
                double v = val[c * vs + m];
                buf[(il * s1 + jl) * ncomp + c] += wll * v;
                buf[(ir * s1 + jl) * ncomp + c] += wlr * v;
                buf[(il * s1 + jr) * ncomp + c] += wul * v;
                buf[(ir * s1 + jr) * ncomp + c] += wur * v;
            }
            continue;
        }
        if (kind == KIND_LINEAR && ndim == 3) {
            int64_t lo[3], hi[3];
            double f[3];
            for (int d = 0; d < 3; ++d) {
                double r = (pos[d * ps + m] - g->lower[d]) * g->inverse[d];
                if (g->code[d] == 1) r = periodic_position(r, g->num[d]);
                lo[d] = left_node(r, g->num[d], g->code[d]);
                hi[d] = right_node(lo[d], g->num[d], g->code[d]);
                f[d] = cell_fraction(r, lo[d]);
            }
            double gx = 1.0 - f[0], gy = 1.0 - f[1], gz = 1.0 - f[2];
            double w00 = gx * gy, w01 = gx * f[1], w10 = f[0] * gy, w11 = f[0] * f[1];
            int64_t r00 = (lo[0] * g->num[1] + lo[1]) * g->num[2], r01 = (lo[0] * g->num[1] + hi[1]) * g->num[2];
            int64_t r10 = (hi[0] * g->num[1] + lo[1]) * g->num[2], r11 = (hi[0] * g->num[1] + hi[1]) * g->num[2];
            for (int64_t c = 0; c < ncomp; ++c) {
                double v = val[c * vs + m];
                buf[(r00 + lo[2]) * ncomp + c] += w00 * gz * v;
                buf[(r00 + hi[2]) * ncomp + c] += w00 * f[2] * v;
                buf[(r01 + lo[2]) * ncomp + c] += w01 * gz * v;
                buf[(r01 + hi[2]) * ncomp + c] += w01 * f[2] * v;
                buf[(r10 + lo[2]) * ncomp + c] += w10 * gz * v;
                buf[(r10 + hi[2]) * ncomp + c] += w10 * f[2] * v;
                buf[(r11 + lo[2]) * ncomp + c] += w11 * gz * v;
                buf[(r11 + hi[2]) * ncomp + c] += w11 * f[2] * v;
            }
            continue;
        }
        int64_t node[3][4] = {{0}};
        double w[3][4] = {{1.0, 0, 0, 0}, {1.0, 0, 0, 0}, {1.0, 0, 0, 0}};
        int size[3] = {1, 1, 1};
        for (int d = 0; d < ndim; ++d)
            size[d] = axis_stencil(kind, degree, (pos[d * ps + m] - g->lower[d]) * g->inverse[d], g->num[d],
                                   g->code[d], node[d], w[d]);
        for (int a = 0; a < size[0]; ++a)
            for (int b = 0; b < size[1]; ++b) {
                double wab = ndim >= 2 ? w[0][a] * w[1][b] : w[0][a];
                for (int e = 0; e < size[2]; ++e) {
                    double weight = ndim == 3 ? wab * w[2][e] : wab;
                    int64_t at = (node[0][a] * s1 + node[1][b] * s2 + node[2][e]) * ncomp;
                    for (int64_t c = 0; c < ncomp; ++c) buf[at + c] += weight * val[c * vs + m];
                }
            }
    }
}

static int64_t grid_nodes(const grid_t *g) { return g->extent[0] * g->extent[1] * g->extent[2]; }

/* scatter: adds.  The thread policy of scatter_quadratic_3d_pyccel.py:147-218; with guard
 * = 0 the MAX_BUFFER_VALUES bound on the private copies is lifted (a labelled variant of the
 * CPU baseline, not the reference's policy). */
void orc_scatter_policy(int kind, int degree, int ndim, const double *pos, int64_t pos_stride, const double *val,
                        int64_t val_stride, int64_t n, int64_t ncomp, double *field, const double *lower,
                        const double *cell, const int64_t *num, const int64_t *code, int threads, int guard) {
    grid_t g;
    make_grid(&g, ndim, kind, degree, lower, cell, num, code);
    const int64_t values = grid_nodes(&g) * ncomp;
    if (threads < 1) threads = 1;
    int serial = threads == 1 || n < MIN_PARALLEL_SCATTER || n < values;
    if (guard && kind != KIND_LINEAR && (int64_t)threads * values > MAX_BUFFER_VALUES) serial = 1;
    if (kind == KIND_LINEAR && ndim == 1 && n < g.extent[0]) serial = 1;
    if (serial) {
        scatter_chunk(kind, degree, &g, pos, pos_stride, val, val_stride, ncomp, field, 0, n);
        return;
    }
    double *buffer = calloc((size_t)threads * (size_t)values, sizeof(double));
    if (!buffer) {
        scatter_chunk(kind, degree, &g, pos, pos_stride, val, val_stride, ncomp, field, 0, n);
        return;
    }
    const int64_t chunk = (n + threads - 1) / threads;
#pragma omp parallel num_threads(threads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t first = (int64_t)t * chunk < n ? (int64_t)t * chunk : n;
        int64_t last = first + chunk < n ? first + chunk : n;
        scatter_chunk(kind, degree, &g, pos, pos_stride, val, val_stride, ncomp, buffer + (size_t)t * values, first,
                      last);
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < values; ++i) {
            double total = 0.0;
            for (int k = 0; k < threads; ++k) total += buffer[(size_t)k * values + i];
            field[i] += total;
        }
    }
    free(buffer);
}

void orc_scatter(int kind, int degree, int ndim, const double *pos, int64_t pos_stride, const double *val,
                 int64_t val_stride, int64_t n, int64_t ncomp, double *field, const double *lower, const double *cell,
                 const int64_t *num, const int64_t *code, int threads) {
    orc_scatter_policy(kind, degree, ndim, pos, pos_stride, val, val_stride, n, ncomp, field, lower, cell, num, code,
                       threads, 1);
}

/* The gather of markers [first, last).  gather_linear_{1,2,3}d_pyccel.py, gather_quadratic_3d_pyccel.py:81-114,
 * gather_spline_3d_pyccel.py:55-90: overwrites result. */
static void gather_chunk(int kind, int degree, const grid_t *g, const double *pos, int64_t ps, const double *field,
                         int64_t ncomp, double *res, int64_t rs, int64_t first, int64_t last) {
    const int ndim = g->ndim;
    const int64_t s1 = g->extent[1] * g->extent[2], s2 = g->extent[2];
    for (int64_t m = first; m < last; ++m) {
        if (kind == KIND_LINEAR && ndim == 1) {
            double rel = (pos[m] - g->lower[0]) * g->inverse[0];
            if (g->code[0] == 1) rel = periodic_position(rel, g->num[0]);
            int64_t l = left_node(rel, g->num[0], g->code[0]);
            int64_t r = right_node(l, g->num[0], g->code[0]);
            double f = cell_fraction(rel, l);
            for (int64_t c = 0; c < ncomp; ++c)
                res[c * rs + m] = (1.0 - f) * field[l * ncomp + c] + f * field[r * ncomp + c];
            continue;
        }
        if (kind == KIND_LINEAR && ndim == 2) {
            double rx = (pos[m] - g->lower[0]) * g->inverse[0];
            if (g->code[0] == 1) rx = periodic_position(rx, g->num[0]);
            int64_t il = left_node(rx, g->num[0], g->code[0]);
            int64_t ir = right_node(il, g->num[0], g->code[0]);
            double fx = cell_fraction(rx, il);
            double ry = (pos[ps + m] - g->lower[1]) * g->inverse[1];
            if (g->code[1] == 1) ry = periodic_position(ry, g->num[1]);
            int64_t jl = left_node(ry, g->num[1], g->code[1]);
            int64_t jr = right_node(jl, g->num[1], g->code[1]);
            double fy = cell_fraction(ry, jl);
            double wll = (1.0 - fx) * (1.0 - fy), wlr = fx * (1.0 - fy), wul = (1.0 - fx) * fy, wur = fx * fy;
            for (int64_t c = 0; c < ncomp; ++c)
                res[c * rs + m] = wll * field[(il * s1 + jl) * ncomp + c] + wlr * field[(ir * s1 + jl) * ncomp + c] +
                                  wul * field[(il * s1 + jr) * ncomp + c] + wur * field[(ir * s1 + jr) * ncomp + c];
            continue;
        }
        if (kind == KIND_LINEAR && ndim == 3) {
            int64_t lo[3], hi[3];
            double f[3];
            for (int d = 0; d < 3; ++d) {
                double r = (pos[d * ps + m] - g->lower[d]) * g->inverse[d];
                if (g->code[d] == 1) r = periodic_position(r, g->num[d]);
                lo[d] = left_node(r, g->num[d], g->code[d]);
                hi[d] = right_node(lo[d], g->num[d], g->code[d]);
                f[d] = cell_fraction(r, lo[d]);
            }
            double gx = 1.0 - f[0], gy = 1.0 - f[1], gz = 1.0 - f[2];
            double w00 = gx * gy, w01 = gx * f[1], w10 = f[0] * gy, w11 = f[0] * f[1];
            int64_t r00 = (lo[0] * g->num[1] + lo[1]) * g->num[2], r01 = (lo[0] * g->num[1] + hi[1]) * g->num[2];
            int64_t r10 = (hi[0] * g->num[1] + lo[1]) * g->num[2], r11 = (hi[0] * g->num[1] + hi[1]) * g->num[2];
            for (int64_t c = 0; c < ncomp; ++c) {
                double t00 = gz * field[(r00 + lo[2]) * ncomp + c] + f[2] * field[(r00 + hi[2]) * ncomp + c];
                double t01 = gz * field[(r01 + lo[2]) * ncomp + c] + f[2] * field[(r01 + hi[2]) * ncomp + c];
                double t10 = gz * field[(r10 + lo[2]) * ncomp + c] + f[2] * field[(r10 + hi[2]) * ncomp + c];
                double t11 = gz * field[(r11 + lo[2]) * ncomp + c] + f[2] * field[(r11 + hi[2]) * ncomp + c];
                res[c * rs + m] = w00 * t00 + w01 * t01 + w10 * t10 + w11 * t11;
            }
            continue;
        }
        int64_t node[3][4] = {{0}};
        double w[3][4] = {{1.0, 0, 0, 0}, {1.0, 0, 0, 0}, {1.0, 0, 0, 0}};
        int size[3] = {1, 1, 1};
        for (int d = 0; d < ndim; ++d)
            size[d] = axis_stencil(kind, degree, (pos[d * ps + m] - g->lower[d]) * g->inverse[d], g->num[d],
                                   g->code[d], node[d], w[d]);
        for (int64_t c = 0; c < ncomp; ++c) {
            double total = 0.0;
            for (int a = 0; a < size[0]; ++a)
                for (int b = 0; b < size[1]; ++b) {
                    double wab = ndim >= 2 ? w[0][a] * w[1][b] : w[0][a];
                    for (int e = 0; e < size[2]; ++e) {
                        double weight = ndim == 3 ? wab * w[2][e] : wab;
                        total += weight * field[(node[0][a] * s1 + node[1][b] * s2 + node[2][e]) * ncomp + c];
                    }
                }
            res[c * rs + m] = total;
        }
    }
}

void orc_gather(int kind, int degree, int ndim, const double *pos, int64_t pos_stride, int64_t n,
                const double *field, int64_t ncomp, double *result, int64_t result_stride, const double *lower,
                const double *cell, const int64_t *num, const int64_t *code, int threads) {
    grid_t g;
    make_grid(&g, ndim, kind, degree, lower, cell, num, code);
    if (threads < 1 || n < MIN_PARALLEL_GATHER) threads = 1;
    const int64_t chunk = (n + threads - 1) / threads;
#pragma omp parallel num_threads(threads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t first = (int64_t)t * chunk < n ? (int64_t)t * chunk : n;
        int64_t last = first + chunk < n ? first + chunk : n;
        gather_chunk(kind, degree, &g, pos, pos_stride, field, ncomp, result, result_stride, first, last);
    }
}

/* ------------------------------------------------------------------------- */
/* Whitney path deposit                                                       */
/* ------------------------------------------------------------------------- */

/* whitney_cell_pyccel.py:52-59 clamp_to_axis */
static inline double clamp_to_axis(double rel, int64_t num, int64_t periodic) {
    if (periodic == 1) return rel;
    return rel < 0.0 ? 0.0 : (rel > (double)num - 1.0 ? (double)num - 1.0 : rel);
}

/* whitney_cell_pyccel.py:62-68 plane_time */
static inline double plane_time(double plane, double rel, double delta) {
    return delta != 0.0 ? (plane - rel) / delta : 2.0;
}

/* whitney_cell_pyccel.py:31-49 edge_path_weight */
static inline double edge_path_weight(double fu, double fw, double du, double dw, int uu, int uw) {
    return (uu == 1 ? fu : 1.0 - fu) * (uw == 1 ? fw : 1.0 - fw) + (uu == uw ? du * dw / 12.0 : -du * dw / 12.0);
}

static inline int finite_all(const double *v, int k) {
    for (int i = 0; i < k; ++i)
        if (!(v[i] > -1.0e300 && v[i] < 1.0e300)) return 0;
    return 1;
}

/* deposit_paths_{1,2,3}d_pyccel.py _deposit_chunk_*: cut at every grid plane, each piece on its cell's edges. */
static void whitney_chunk(const grid_t *g, const double *st, int64_t ss, const double *en, int64_t es,
                          const double *q, double *edges, int64_t first, int64_t last) {
    const int ndim = g->ndim;
    const int64_t s1 = g->num[1] * g->num[2], s2 = g->num[2];
    for (int64_t p = first; p < last; ++p) {
        double charge = q[p];
        if (charge == 0.0) continue;
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, ab[6];
        for (int d = 0; d < ndim; ++d) {
            a[d] = (st[d * ss + p] - g->lower[d]) * g->inverse[d];
            b[d] = (en[d * es + p] - g->lower[d]) * g->inverse[d];
        }
        for (int d = 0; d < ndim; ++d) {
            ab[2 * d] = a[d];
            ab[2 * d + 1] = b[d];
        }
        if (!finite_all(ab, 2 * ndim)) continue;
        double delta[3] = {0, 0, 0}, step[3] = {0, 0, 0}, plane[3] = {0, 0, 0}, x0[3] = {0, 0, 0}, x1[3] = {0, 0, 0};
        for (int d = 0; d < ndim; ++d) {
            delta[d] = b[d] - a[d];
            step[d] = 1.0;
            plane[d] = floor(a[d]) + 1.0;
            if (delta[d] < 0.0) {
                step[d] = -1.0;
                plane[d] = ceil(a[d]) - 1.0;
            }
            x0[d] = clamp_to_axis(a[d], g->num[d], g->code[d]);
        }
        double time = 0.0;
        while (time < 1.0) {
            double following = 1.0, times[3] = {2.0, 2.0, 2.0};
            if (ndim == 1) {
                /* deposit_paths_1d_pyccel.py:65-73 */
                following = plane_time(plane[0], a[0], delta[0]);
                if (following < 1.0) {
                    x1[0] = clamp_to_axis(a[0] + following * delta[0], g->num[0], g->code[0]);
                    plane[0] += step[0];
                } else {
                    following = 1.0;
                    x1[0] = clamp_to_axis(b[0], g->num[0], g->code[0]);
                }
            } else {
                for (int d = 0; d < ndim; ++d) {
                    times[d] = plane_time(plane[d], a[d], delta[d]);
                    if (times[d] < following) following = times[d];
                }
                for (int d = 0; d < ndim; ++d)
                    x1[d] = following < 1.0 ? clamp_to_axis(a[d] + following * delta[d], g->num[d], g->code[d])
                                            : clamp_to_axis(b[d], g->num[d], g->code[d]);
                for (int d = 0; d < ndim; ++d)
                    if (times[d] <= following) plane[d] += step[d];
            }
            int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            double f[3] = {0, 0, 0}, dd[3] = {0, 0, 0}, along[3] = {0, 0, 0};
            for (int d = 0; d < ndim; ++d) {
                double mid = 0.5 * (x0[d] + x1[d]);
                if (g->code[d] == 1) mid = periodic_position(mid, g->num[d]);
                lo[d] = left_node(mid, g->num[d], g->code[d]);
                hi[d] = right_node(lo[d], g->num[d], g->code[d]);
                f[d] = cell_fraction(mid, lo[d]);
                dd[d] = x1[d] - x0[d];
                along[d] = charge * dd[d];
            }
            if (ndim == 1) {
                edges[lo[0]] += charge * (x1[0] - x0[0]);
            } else if (ndim == 2) {
                /* deposit_paths_2d_pyccel.py:130-135 */
                edges[(lo[0] * s1 + lo[1]) * 2 + 0] += (1.0 - f[1]) * along[0];
                edges[(lo[0] * s1 + hi[1]) * 2 + 0] += f[1] * along[0];
                edges[(lo[0] * s1 + lo[1]) * 2 + 1] += (1.0 - f[0]) * along[1];
                edges[(hi[0] * s1 + lo[1]) * 2 + 1] += f[0] * along[1];
            } else {
                /* deposit_paths_3d_pyccel.py:176-212 */
#define E3(i, j, k, a) edges[((i)*s1 + (j)*s2 + (k)) * 3 + (a)]
                E3(lo[0], lo[1], lo[2], 0) += along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 0, 0);
                E3(lo[0], hi[1], lo[2], 0) += along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 1, 0);
                E3(lo[0], lo[1], hi[2], 0) += along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 0, 1);
                E3(lo[0], hi[1], hi[2], 0) += along[0] * edge_path_weight(f[1], f[2], dd[1], dd[2], 1, 1);
                E3(lo[0], lo[1], lo[2], 1) += along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 0, 0);
                E3(hi[0], lo[1], lo[2], 1) += along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 1, 0);
                E3(lo[0], lo[1], hi[2], 1) += along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 0, 1);
                E3(hi[0], lo[1], hi[2], 1) += along[1] * edge_path_weight(f[0], f[2], dd[0], dd[2], 1, 1);
                E3(lo[0], lo[1], lo[2], 2) += along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 0, 0);
                E3(hi[0], lo[1], lo[2], 2) += along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 1, 0);
                E3(lo[0], hi[1], lo[2], 2) += along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 0, 1);
                E3(hi[0], hi[1], lo[2], 2) += along[2] * edge_path_weight(f[0], f[1], dd[0], dd[1], 1, 1);
#undef E3
            }
            time = following;
            for (int d = 0; d < ndim; ++d) x0[d] = x1[d];
        }
    }
}

/* deposit_paths_2d_pyccel.py:141-183: per-thread copies of the edges, added thread by thread. */
void orc_deposit_whitney(int ndim, const double *start, int64_t start_stride, const double *end, int64_t end_stride,
                         const double *q, int64_t n, double *edges, const double *lower, const double *cell,
                         const int64_t *num, const int64_t *code, int threads) {
    grid_t g;
    make_grid(&g, ndim, KIND_LINEAR, 1, lower, cell, num, code);
    const int64_t nodes = grid_nodes(&g), values = nodes * ndim;
    if (threads < 1 || n < MIN_PARALLEL_SCATTER || n < nodes) threads = 1;
    double *buffer = calloc((size_t)threads * (size_t)values, sizeof(double));
    if (!buffer) return;
    const int64_t chunk = (n + threads - 1) / threads;
#pragma omp parallel num_threads(threads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t first = (int64_t)t * chunk < n ? (int64_t)t * chunk : n;
        int64_t last = first + chunk < n ? first + chunk : n;
        whitney_chunk(&g, start, start_stride, end, end_stride, q, buffer + (size_t)t * values, first, last);
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < values; ++i)
            for (int k = 0; k < threads; ++k) edges[i] += buffer[(size_t)k * values + i];
    }
    free(buffer);
}

/* ------------------------------------------------------------------------- */
/* Spline path deposit                                                        */
/* ------------------------------------------------------------------------- */

/* spline_edges_pyccel.py:19-24 floor_int */
static inline int64_t floor_int(double x) {
    int64_t whole = (int64_t)x;
    if (x < (double)whole) whole -= 1;
    return whole;
}

/* spline_edges_pyccel.py:86-138 along_stencil */
static void along_stencil(double rel, int64_t num, int64_t periodic, int degree, int64_t *edge, double *weight,
                          int *exists) {
    if (periodic == 2) {
        clamped_cell(rel, num, degree, 1, edge, weight);
        for (int a = 0; a < degree; ++a) exists[a] = 1;
        return;
    }
    int lower = degree - 1;
    double r = rel;
    if (periodic == 1 && !(r >= 0.0 && r < (double)num)) r = r - (double)num * floor(r / (double)num);
    double shifted = r - 0.5;
    int64_t middle;
    if (lower % 2 == 0) {
        middle = floor_int(shifted + 0.5);
        if (periodic == 0 && middle > num - 2) middle = num - 2;
    } else {
        middle = floor_int(shifted);
    }
    double offset = shifted - (double)middle;
    if (lower == 0) {
        weight[0] = 1.0;
    } else if (lower == 1) {
        weight[0] = 1.0 - offset;
        weight[1] = offset;
    } else {
        weight[0] = 0.5 * (0.5 - offset) * (0.5 - offset);
        weight[1] = 0.75 - offset * offset;
        weight[2] = 0.5 * (0.5 + offset) * (0.5 + offset);
    }
    int64_t first = middle - lower / 2;
    for (int a = 0; a < degree; ++a) {
        int64_t k = first + a;
        if (periodic == 1) {
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

/* spline_edges_pyccel.py:141-177 gauss_rule */
static void gauss_rule(int count, double *point, double *weight) {
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

/* spline_edges_pyccel.py:180-230 path_cuts */
static int path_cuts(const double *r0, const double *r1, int ndim, double shift, const int64_t *num,
                     const int64_t *periodic, double *cuts) {
    int count = 1;
    cuts[0] = 0.0;
    for (int d = 0; d < ndim; ++d) {
        double delta = r1[d] - r0[d];
        if (delta == 0.0) continue;
        double low = delta > 0.0 ? r0[d] : r1[d];
        double high = delta > 0.0 ? r1[d] : r0[d];
        double offset = periodic[d] == 2 ? 0.0 : shift;
        double plane = (double)floor_int(low - offset) + 1.0 + offset;
        while (plane < high) {
            double t = (plane - r0[d]) / delta;
            if (t > 0.0 && t < 1.0 && count < 31) cuts[count++] = t;
            plane += 1.0;
        }
        if (periodic[d] != 1) {
            double t = (0.0 - r0[d]) / delta;
            if (t > 0.0 && t < 1.0 && count < 31) cuts[count++] = t;
            t = ((double)num[d] - 1.0 - r0[d]) / delta;
            if (t > 0.0 && t < 1.0 && count < 31) cuts[count++] = t;
        }
    }
    cuts[count++] = 1.0;
    for (int i = 2; i < count - 1; ++i) {
        double value = cuts[i];
        int j = i - 1;
        while (j > 0 && cuts[j] > value) {
            cuts[j + 1] = cuts[j];
            --j;
        }
        cuts[j + 1] = value;
    }
    return count;
}

/* deposit_paths_spline_{1,2,3}d_pyccel.py: stretches of at most 4 cells, cut at knot planes and walls,
 * Gauss-Legendre per piece; edges num (+ degree - 1 on a clamped axis) per axis, times ndim. */
void orc_deposit_spline(int ndim, int degree, const double *start, int64_t start_stride, const double *end,
                        int64_t end_stride, const double *q, int64_t n, double *edges, const double *lower,
                        const double *cell, const int64_t *num, const int64_t *code) {
    grid_t g;
    make_grid(&g, ndim, KIND_SPLINE, degree, lower, cell, num, code);
    const int64_t s1 = g.extent[1] * g.extent[2], s2 = g.extent[2];
    double shift = degree % 2 == 0 ? 0.5 : 0.0;
    int num_points = (degree - 1 + (ndim - 1) * degree) / 2 + 1;
    double point[5], gw[5], cuts[32];
    gauss_rule(num_points, point, gw);
    for (int64_t p = 0; p < n; ++p) {
        double charge = q[p];
        if (charge == 0.0) continue;
        int finite = 1;
        for (int d = 0; d < ndim; ++d) {
            if (!(start[d * start_stride + p] - start[d * start_stride + p] == 0.0)) finite = 0;
            if (!(end[d * end_stride + p] - end[d * end_stride + p] == 0.0)) finite = 0;
        }
        if (!finite) continue;
        double r0[3], r1[3], extent = 0.0;
        for (int d = 0; d < ndim; ++d) {
            r0[d] = (start[d * start_stride + p] - g.lower[d]) * g.inverse[d];
            r1[d] = (end[d * end_stride + p] - g.lower[d]) * g.inverse[d];
            double span = r1[d] > r0[d] ? r1[d] - r0[d] : r0[d] - r1[d];
            if (span > extent) extent = span;
        }
        int64_t stretches = (int64_t)(extent / 4.0) + 1;
        for (int64_t s = 0; s < stretches; ++s) {
            double a0[3], a1[3];
            for (int d = 0; d < ndim; ++d) {
                a0[d] = r0[d] + (r1[d] - r0[d]) * (double)s / (double)stretches;
                a1[d] = r0[d] + (r1[d] - r0[d]) * (double)(s + 1) / (double)stretches;
            }
            int num_cuts = path_cuts(a0, a1, ndim, shift, g.num, g.code, cuts);
            for (int piece = 0; piece < num_cuts - 1; ++piece) {
                double t0 = cuts[piece], t1 = cuts[piece + 1];
                if (t1 <= t0) continue;
                double p0[3], stp[3];
                for (int d = 0; d < ndim; ++d) {
                    double begin = a0[d] + t0 * (a1[d] - a0[d]);
                    double finish = a0[d] + t1 * (a1[d] - a0[d]);
                    if (g.code[d] != 1) {
                        if (begin < 0.0) begin = 0.0;
                        if (begin > (double)(g.num[d] - 1)) begin = (double)g.num[d] - 1.0;
                        if (finish < 0.0) finish = 0.0;
                        if (finish > (double)(g.num[d] - 1)) finish = (double)g.num[d] - 1.0;
                    }
                    p0[d] = begin;
                    stp[d] = finish - begin;
                }
                for (int gp = 0; gp < num_points; ++gp) {
                    double rel[3];
                    for (int d = 0; d < ndim; ++d) rel[d] = p0[d] + point[gp] * stp[d];
                    for (int axis = 0; axis < ndim; ++axis) {
                        double amount = charge * gw[gp] * stp[axis];
                        if (amount == 0.0) continue;
                        int64_t idx[3][4] = {{0}};
                        double w[3][4] = {{1.0, 0, 0, 0}, {1.0, 0, 0, 0}, {1.0, 0, 0, 0}};
                        int ex[4] = {1, 1, 1, 1}, size[3] = {1, 1, 1};
                        for (int d = 0; d < ndim; ++d) {
                            if (d == axis) {
                                along_stencil(rel[d], g.num[d], g.code[d], degree, idx[d], w[d], ex);
                                size[d] = degree;
                            } else {
                                spline_cell(rel[d], g.num[d], g.code[d], degree, idx[d], w[d]);
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
                                    if (ndim >= 2) v = v * w[1][b];
                                    if (ndim == 3) v = v * w[2][e];
                                    edges[(idx[0][a] * s1 + idx[1][b] * s2 + idx[2][e]) * ndim + axis] += v;
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Edge gathers (the transpose of the path deposits)                          */
/* ------------------------------------------------------------------------- */

/* gather_edges_{1,2,3}d_pyccel.py (1D :27-35, 2D :37-63, 3D :40-86): the Whitney field of
 * the edge line integrals at each marker, in LINEAR's cell (wrapped on a periodic axis,
 * clamped on a bounded one): along axis a the edges of the cell along a, weighted by the
 * linear weights edge_weight(f, 0/1) = 1 - f, f across the cell (whitney_cell_pyccel.py:
 * 23-27), summed in the reference's order, divided by the cell size along a. */
void orc_gather_edges_whitney(int ndim, const double *pos, int64_t ps, int64_t n, const double *edges,
                              double *result, int64_t rs, const double *lower, const double *cell,
                              const int64_t *num, const int64_t *code) {
    grid_t g;
    make_grid(&g, ndim, KIND_LINEAR, 1, lower, cell, num, code);
    const int64_t s1 = g.extent[1] * g.extent[2], s2 = g.extent[2];
    for (int64_t m = 0; m < n; ++m) {
        int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        double f[3] = {0, 0, 0};
        for (int d = 0; d < ndim; ++d) {
            double rel = (pos[d * ps + m] - g.lower[d]) * g.inverse[d];
            if (g.code[d] == 1) rel = periodic_position(rel, g.num[d]);
            lo[d] = left_node(rel, g.num[d], g.code[d]);
            hi[d] = right_node(lo[d], g.num[d], g.code[d]);
            f[d] = cell_fraction(rel, lo[d]);
        }
#define E(i, j, k, a) edges[((i) * (ndim == 3 ? s1 : (ndim == 2 ? g.extent[1] : 1)) + (ndim == 3 ? (j) * s2 : (ndim == 2 ? (j) : 0)) + (ndim == 3 ? (k) : 0)) * ndim + (a)]
        if (ndim == 1) {
            result[m] = edges[lo[0]] * g.inverse[0];
        } else if (ndim == 2) {
            result[m] = ((1.0 - f[1]) * E(lo[0], lo[1], 0, 0) + f[1] * E(lo[0], hi[1], 0, 0)) * g.inverse[0];
            result[rs + m] = ((1.0 - f[0]) * E(lo[0], lo[1], 0, 1) + f[0] * E(hi[0], lo[1], 0, 1)) * g.inverse[1];
        } else {
            const double wy0 = 1.0 - f[1], wy1 = f[1], wz0 = 1.0 - f[2], wz1 = f[2], wx0 = 1.0 - f[0], wx1 = f[0];
            result[m] = (wy0 * wz0 * E(lo[0], lo[1], lo[2], 0) + wy1 * wz0 * E(lo[0], hi[1], lo[2], 0) +
                         wy0 * wz1 * E(lo[0], lo[1], hi[2], 0) + wy1 * wz1 * E(lo[0], hi[1], hi[2], 0)) * g.inverse[0];
            result[rs + m] = (wx0 * wz0 * E(lo[0], lo[1], lo[2], 1) + wx1 * wz0 * E(hi[0], lo[1], lo[2], 1) +
                              wx0 * wz1 * E(lo[0], lo[1], hi[2], 1) + wx1 * wz1 * E(hi[0], lo[1], hi[2], 1)) * g.inverse[1];
            result[2 * rs + m] = (wx0 * wy0 * E(lo[0], lo[1], lo[2], 2) + wx1 * wy0 * E(hi[0], lo[1], lo[2], 2) +
                                  wx0 * wy1 * E(lo[0], hi[1], lo[2], 2) + wx1 * wy1 * E(hi[0], hi[1], lo[2], 2)) * g.inverse[2];
        }
#undef E
    }
}

/* gather_edges_spline_{1,2,3}d_pyccel.py (2D :52-82, 3D :70-130): the field of the spline
 * edge functions B_{p-1}(r_a - k_a - 1/2) prod_{b != a} B_p(r_b - k_b) at each marker,
 * positions clamped to [0, num - 1] on a non-periodic axis; along axis a the along_stencil
 * edges that exist times the spline_cell nodes across, the products left to right and the
 * sums in the reference's loop order, divided by the cell size along a. */
void orc_gather_edges_spline(int ndim, int degree, const double *pos, int64_t ps, int64_t n, const double *edges,
                             double *result, int64_t rs, const double *lower, const double *cell,
                             const int64_t *num, const int64_t *code) {
    grid_t g;
    make_grid(&g, ndim, KIND_SPLINE, degree, lower, cell, num, code);
    const int64_t s0 = ndim == 3 ? g.extent[1] * g.extent[2] : (ndim == 2 ? g.extent[1] : 1);
    const int64_t s1 = ndim == 3 ? g.extent[2] : 1;
    for (int64_t m = 0; m < n; ++m) {
        double rel[3] = {0, 0, 0};
        for (int d = 0; d < ndim; ++d) {
            double r = (pos[d * ps + m] - g.lower[d]) * g.inverse[d];
            if (g.code[d] != 1) {
                if (r < 0.0) r = 0.0;
                if (r > (double)(g.num[d] - 1)) r = (double)g.num[d] - 1.0;
            }
            rel[d] = r;
        }
        for (int axis = 0; axis < ndim; ++axis) {
            int64_t idx[3][4] = {{0}, {0}, {0}};
            double w[3][4] = {{1.0}, {1.0}, {1.0}};
            int ex[4] = {1, 1, 1, 1}, size[3] = {1, 1, 1};
            for (int d = 0; d < ndim; ++d) {
                if (d == axis) {
                    along_stencil(rel[d], g.num[d], g.code[d], degree, idx[d], w[d], ex);
                    size[d] = degree;
                } else {
                    spline_cell(rel[d], g.num[d], g.code[d], degree, idx[d], w[d]);
                    size[d] = degree + 1;
                }
            }
            double total = 0.0;
            for (int a = 0; a < size[0]; ++a) {
                if (axis == 0 && !ex[a]) continue;
                for (int b = 0; b < size[1]; ++b) {
                    if (axis == 1 && !ex[b]) continue;
                    for (int e = 0; e < size[2]; ++e) {
                        if (axis == 2 && !ex[e]) continue;
                        double v = w[0][a];
                        if (ndim >= 2) v = v * w[1][b];
                        if (ndim == 3) v = v * w[2][e];
                        total += v * edges[(idx[0][a] * s0 + idx[1][b] * s1 + idx[2][e]) * ndim + axis];
                    }
                }
            }
            result[axis * rs + m] = total * g.inverse[axis];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Forms of the spline de Rham sequence (faces, cells)                        */
/* ------------------------------------------------------------------------- */

/* spline_edges_pyccel.py:233-245 staggered: edges along their axis, faces across their
 * normal, the top form along every axis. */
static int form_staggered(int form, int ndim, int component, int axis) {
    if (form == ndim) return 1;
    if (form == 1) return axis == component;
    if (form == 2 && ndim == 3) return axis != component;
    return 0;
}

/* spline_forms.py:39-43 num_components: ndim choose form. */
int orc_form_components(int form, int ndim) {
    if (form < 1 || form > ndim || ndim > 3) return 0;
    return form == ndim ? 1 : ndim;
}

/* One axis of component c: along_stencil (degree edges, some absent on a bounded axis)
 * where c is staggered, spline_cell (degree + 1 nodes, all present) elsewhere. */
static int form_axis(int form, int ndim, int c, int d, int degree, double rel, const grid_t *g, int64_t *idx,
                     double *w, int *ex) {
    if (form_staggered(form, ndim, c, d)) {
        along_stencil(rel, g->num[d], g->code[d], degree, idx, w, ex);
        return degree;
    }
    spline_cell(rel, g->num[d], g->code[d], degree, idx, w);
    for (int a = 0; a <= degree; ++a) ex[a] = 1;
    return degree + 1;
}

/* The marker's position in cells, clamped to [0, num - 1] on a non-periodic axis
 * (gather_forms_spline_2d_pyccel.py:52-63). */
static void form_relative(const grid_t *g, const double *pos, int64_t ps, int64_t m, double rel[3]) {
    rel[0] = rel[1] = rel[2] = 0.0;
    for (int d = 0; d < g->ndim; ++d) {
        double r = (pos[d * ps + m] - g->lower[d]) * g->inverse[d];
        if (g->code[d] != 1) {
            if (r < 0.0) r = 0.0;
            if (r > (double)(g->num[d] - 1)) r = (double)g->num[d] - 1.0;
        }
        rel[d] = r;
    }
}

/* gather_forms_spline_{1,2,3}d_pyccel.py _gather_chunk (2D :25-90, 3D :30-130): per
 * component the sum over the present k-cells of the weight product (left to right) times
 * the field, loops x outer to z inner; overwrites result (ncomp, n). */
static void gather_forms_chunk(int degree, int form, const grid_t *gp, const double *pos, int64_t ps,
                               const double *field, double *result, int64_t rs, int64_t first, int64_t last) {
    const grid_t g = *gp;
    const int ndim = g.ndim;
    const int ncomp = orc_form_components(form, ndim);
    const int64_t s0 = g.extent[1] * g.extent[2], s1 = g.extent[2];
    for (int64_t m = first; m < last; ++m) {
        double rel[3];
        form_relative(&g, pos, ps, m, rel);
        for (int c = 0; c < ncomp; ++c) {
            int64_t idx[3][4] = {{0}, {0}, {0}};
            double w[3][4] = {{1.0}, {1.0}, {1.0}};
            int ex[3][4] = {{1, 1, 1, 1}, {1, 1, 1, 1}, {1, 1, 1, 1}}, size[3] = {1, 1, 1};
            for (int d = 0; d < ndim; ++d) size[d] = form_axis(form, ndim, c, d, degree, rel[d], &g, idx[d], w[d], ex[d]);
            double total = 0.0;
            for (int a = 0; a < size[0]; ++a) {
                if (!ex[0][a]) continue;
                for (int b = 0; b < size[1]; ++b) {
                    if (!ex[1][b]) continue;
                    for (int e = 0; e < size[2]; ++e) {
                        if (!ex[2][e]) continue;
                        double v = w[0][a];
                        if (ndim >= 2) v = v * w[1][b];
                        if (ndim == 3) v = v * w[2][e];
                        total += v * field[(idx[0][a] * s0 + idx[1][b] * s1 + idx[2][e]) * ncomp + c];
                    }
                }
            }
            result[c * rs + m] = total;
        }
    }
}

/* gather_forms_spline_2d_pyccel.py:93-150: serial for one thread or fewer than 4096 markers,
 * else contiguous chunks of the markers, one per thread. */
void orc_gather_forms(int ndim, int degree, int form, const double *pos, int64_t ps, int64_t n, const double *field,
                      double *result, int64_t rs, const double *lower, const double *cell, const int64_t *num,
                      const int64_t *code, int threads) {
    grid_t g;
    make_grid(&g, ndim, KIND_SPLINE, degree, lower, cell, num, code);
    if (threads <= 1 || n < MIN_PARALLEL_GATHER) {
        gather_forms_chunk(degree, form, &g, pos, ps, field, result, rs, 0, n);
        return;
    }
    const int64_t chunk = (n + threads - 1) / threads;
#pragma omp parallel num_threads(threads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t first = (int64_t)t * chunk < n ? (int64_t)t * chunk : n;
        int64_t last = first + chunk < n ? first + chunk : n;
        gather_forms_chunk(degree, form, &g, pos, ps, field, result, rs, first, last);
    }
}

/* scatter_forms_spline_{1,2,3}d_pyccel.py _scatter_chunk (3D :23-121): a marker whose values
 * are all zero is skipped before its position is read; each present k-cell adds
 * w_x w_y w_z value. */
static void forms_scatter_chunk(int degree, int form, const grid_t *g, const double *pos, int64_t ps,
                                const double *val, int64_t vs, int ncomp, double *buf, int64_t first, int64_t last) {
    const int ndim = g->ndim;
    const int64_t s0 = g->extent[1] * g->extent[2], s1 = g->extent[2];
    for (int64_t m = first; m < last; ++m) {
        int live = 0;
        for (int c = 0; c < ncomp; ++c)
            if (val[c * vs + m] != 0.0) live = 1;
        if (!live) continue;
        double rel[3];
        form_relative(g, pos, ps, m, rel);
        for (int c = 0; c < ncomp; ++c) {
            const double value = val[c * vs + m];
            int64_t idx[3][4] = {{0}, {0}, {0}};
            double w[3][4] = {{1.0}, {1.0}, {1.0}};
            int ex[3][4] = {{1, 1, 1, 1}, {1, 1, 1, 1}, {1, 1, 1, 1}}, size[3] = {1, 1, 1};
            for (int d = 0; d < ndim; ++d) size[d] = form_axis(form, ndim, c, d, degree, rel[d], g, idx[d], w[d], ex[d]);
            for (int a = 0; a < size[0]; ++a) {
                if (!ex[0][a]) continue;
                for (int b = 0; b < size[1]; ++b) {
                    if (!ex[1][b]) continue;
                    for (int e = 0; e < size[2]; ++e) {
                        if (!ex[2][e]) continue;
                        double v = w[0][a];
                        if (ndim >= 2) v = v * w[1][b];
                        if (ndim == 3) v = v * w[2][e];
                        buf[(idx[0][a] * s0 + idx[1][b] * s1 + idx[2][e]) * ncomp + c] += v * value;
                    }
                }
            }
        }
    }
}

/* scatter_forms_spline_3d_pyccel.py:150-210: adds; the node deposits' thread policy
 * (serial for few markers or a large grid, else per-thread copies summed in thread order). */
void orc_scatter_forms(int ndim, int degree, int form, const double *pos, int64_t ps, const double *val, int64_t vs,
                       int64_t n, double *field, const double *lower, const double *cell, const int64_t *num,
                       const int64_t *code, int threads) {
    grid_t g;
    make_grid(&g, ndim, KIND_SPLINE, degree, lower, cell, num, code);
    const int ncomp = orc_form_components(form, ndim);
    const int64_t values = grid_nodes(&g) * ncomp;
    if (threads < 1) threads = 1;
    int serial = threads == 1 || n < MIN_PARALLEL_SCATTER || n < values || (int64_t)threads * values > MAX_BUFFER_VALUES;
    double *buffer = serial ? NULL : calloc((size_t)threads * (size_t)values, sizeof(double));
    if (!buffer) {
        forms_scatter_chunk(degree, form, &g, pos, ps, val, vs, ncomp, field, 0, n);
        return;
    }
    const int64_t chunk = (n + threads - 1) / threads;
#pragma omp parallel num_threads(threads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t first = (int64_t)t * chunk < n ? (int64_t)t * chunk : n;
        int64_t last = first + chunk < n ? first + chunk : n;
        forms_scatter_chunk(degree, form, &g, pos, ps, val, vs, ncomp, buffer + (size_t)t * values, first, last);
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < values; ++i) {
            double total = 0.0;
            for (int k = 0; k < threads; ++k) total += buffer[(size_t)k * values + i];
            field[i] += total;
        }
    }
    free(buffer);
}

/* ------------------------------------------------------------------------- */
/* Halo fill and fold, gradient and divergence                                */
/* ------------------------------------------------------------------------- */

/* Storage of shape[0..ndim) nodes (halo h included on every side) times ncomp.  Axis by axis over the full
 * storage of the other axes (halos included), so corners reach the diagonal in two steps:
 *   fill  (mpiarray update_halos, field.py:601-613): on a periodic axis with one block the low halo takes the
 *         last h interior layers and the high halo the first h; a bounded axis is left unchanged.
 *   fold  (mpiarray accumulate_halos, field.py:615-630): on a periodic axis the high halo is added into the first
 *         h interior layers and the low halo into the last h; then every halo layer is cleared (a bounded axis
 *         drops what its halos hold). */
static void halo_axis(int mode, int ndim, const int64_t *shape, int64_t ncomp, int64_t h, int axis, int periodic,
                      double *data) {
    int64_t outer = 1, inner = ncomp;
    for (int d = 0; d < axis; ++d) outer *= shape[d];
    for (int d = axis + 1; d < ndim; ++d) inner *= shape[d];
    const int64_t len = shape[axis], n = len - 2 * h;
    for (int64_t o = 0; o < outer; ++o) {
        double *base = data + o * len * inner;
        for (int64_t l = 0; l < h; ++l) {
            double *lo_halo = base + l * inner, *hi_halo = base + (n + h + l) * inner;
            double *lo_in = base + (h + l) * inner, *hi_in = base + (n + l) * inner;
            for (int64_t i = 0; i < inner; ++i) {
                if (mode == 0) {
                    if (periodic) {
                        lo_halo[i] = hi_in[i];
                        hi_halo[i] = lo_in[i];
                    }
                } else {
                    if (periodic) {
                        lo_in[i] += hi_halo[i];
                        hi_in[i] += lo_halo[i];
                    }
                }
            }
        }
        if (mode == 1)
            for (int64_t l = 0; l < h; ++l)
                for (int64_t i = 0; i < inner; ++i) {
                    base[l * inner + i] = 0.0;
                    base[(n + h + l) * inner + i] = 0.0;
                }
    }
}

void orc_halo_fill(int ndim, const int64_t *shape, int64_t ncomp, int64_t halo, const int64_t *periodic, double *data) {
    for (int d = 0; d < ndim; ++d) halo_axis(0, ndim, shape, ncomp, halo, d, (int)periodic[d], data);
}

void orc_halo_fold(int ndim, const int64_t *shape, int64_t ncomp, int64_t halo, const int64_t *periodic, double *data) {
    for (int d = 0; d < ndim; ++d) halo_axis(1, ndim, shape, ncomp, halo, d, (int)periodic[d], data);
}

/* edges.py:240-259 divergence_block: G^T j per node (periodic roll, bounded drop of the last edge). */
void orc_divergence(int ndim, const int64_t *shape, const int64_t *periodic, const double *edges, double *out) {
    int64_t total = 1;
    for (int d = 0; d < ndim; ++d) total *= shape[d];
    memset(out, 0, (size_t)total * sizeof(double));
    for (int axis = 0; axis < ndim; ++axis) {
        int64_t stride = 1;
        for (int d = axis + 1; d < ndim; ++d) stride *= shape[d];
        for (int64_t i = 0; i < total; ++i) {
            int64_t k = (i / stride) % shape[axis];
            double flow = edges[i * ndim + axis];
            if (periodic[axis]) {
                int64_t prev = k == 0 ? i + (shape[axis] - 1) * stride : i - stride;
                out[i] += edges[prev * ndim + axis] - flow;
            } else if (k < shape[axis] - 1) {
                out[i] -= flow;
                out[i + stride] += flow;
            }
        }
    }
}

/* edges.py:218-237 gradient_block */
void orc_gradient(int ndim, const int64_t *shape, const int64_t *periodic, const double *nodes, double *out) {
    int64_t total = 1;
    for (int d = 0; d < ndim; ++d) total *= shape[d];
    memset(out, 0, (size_t)total * ndim * sizeof(double));
    for (int axis = 0; axis < ndim; ++axis) {
        int64_t stride = 1;
        for (int d = axis + 1; d < ndim; ++d) stride *= shape[d];
        for (int64_t i = 0; i < total; ++i) {
            int64_t k = (i / stride) % shape[axis];
            if (periodic[axis]) {
                int64_t next = k == shape[axis] - 1 ? i - (shape[axis] - 1) * stride : i + stride;
                out[i * ndim + axis] = nodes[next] - nodes[i];
            } else if (k < shape[axis] - 1) {
                out[i * ndim + axis] = nodes[i + stride] - nodes[i];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Stencil probes for the unit tests                                          */
/* ------------------------------------------------------------------------- */

int orc_axis_stencil(int kind, int degree, double rel, int64_t num, int64_t code, int64_t *node, double *weight) {
    return axis_stencil(kind, degree, rel, num, code, node, weight);
}

void orc_along_stencil(double rel, int64_t num, int64_t code, int degree, int64_t *edge, double *weight, int *exists) {
    along_stencil(rel, num, code, degree, edge, weight, exists);
}

int orc_path_cuts(const double *r0, const double *r1, int ndim, double shift, const int64_t *num,
                  const int64_t *code, double *cuts) {
    return path_cuts(r0, r1, ndim, shift, num, code, cuts);
}
