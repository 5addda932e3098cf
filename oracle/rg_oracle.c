/*
 * rg_oracle.c -- CPU restatement of regridding's conservative 2-D hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the shared
 * library built from it, and only as the checker / CPU baseline.  The product
 * (regridding_amd) never links, loads or calls it.
 *
 * Every function follows the reference's algorithm statement by statement so
 * that, compiled with IEEE semantics (-O2 -ffp-contract=off, no fast-math),
 * it produces the same discrete decisions and the same bits as the reference
 * run in exact-order (no fastmath) mode.  Citations are to files under the
 * reference repository (sun-data/regridding):
 *
 *   geometry.py                                   2-D primitives
 *   _interp_ndarray.py                            _bilinear_interpolation
 *   _weights/_weights_conservative_2d/_grids.py   areas, boundary, cell search
 *   _weights/_weights_conservative_2d/_arrays.py  index helpers
 *   _weights/_weights_conservative_2d/_weights_conservative_2d.py  the sweep
 *   _weights/_weights_arrays.py                   coalesce (stable sort + reduceat)
 *   _regrid/_regrid_from_weights.py               apply
 *   _find_indices/_find_indices_{brute,searchsorted}.py   1-D search
 *
 * The row loop of the sweep runs in parallel with OpenMP, as the reference's
 * numba.prange over rows does (_weights_conservative_2d.py:286); per-row
 * triple lists are concatenated in row order afterwards, so the output does
 * not depend on the thread count.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   geometry primitives    reference truth tables (test_geometry.py:112-311,
 *                          432-495) ported as known answers
 *   grid_volume            _grids_test.py:7-42 known answers; exact rational
 *                          shoelace on random quads
 *   index_of_point_*       _grids_test.py:45-122 known answers; brute force
 *   sweep + coalesce       _weights_conservative_2d_test.py:16-301 known
 *                          answers; closed form (rectilinear overlap); brute
 *                          force polygon clipping; conservation invariants
 *   reduceat order         numpy.add.reduceat itself, bit for bit
 *   apply                  numpy.add.at on shuffled triples with repeated
 *                          pairs and mixed magnitudes, bit for bit (the
 *                          sequential storage-order sum); dense matrix
 *                          product; regrid == weights + apply
 *   find_indices 1-D       test_find_indices.py known answers; np.searchsorted
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RGO_MAXSIZE INT64_MAX /* sys.maxsize */

/* ------------------------------------------------------------------------ */
/* geometry.py                                                              */
/* ------------------------------------------------------------------------ */

/* geometry.py:153-284  bounding_boxes_intersect_2d (touching counts) */
static int bbox_intersect(double xp1, double yp1, double xp2, double yp2,
                          double xq1, double yq1, double xq2, double yq2)
{
    double s;
    if (xp1 > xp2) { s = xp1; xp1 = xp2; xp2 = s; }
    if (xq1 > xq2) { s = xq1; xq1 = xq2; xq2 = s; }
    if (!(xp1 <= xq2 && xq1 <= xp2))
        return 0;
    if (yp1 > yp2) { s = yp1; yp1 = yp2; yp2 = s; }
    if (yq1 > yq2) { s = yq1; yq1 = yq2; yq2 = s; }
    if (!(yp1 <= yq2 && yq1 <= yp2))
        return 0;
    return 1;
}

/* geometry.py:370-445  two_line_segment_intersection_parameters */
static void seg_params(double x1, double y1, double x2, double y2,
                       double x3, double y3, double x4, double y4,
                       double *t, double *u)
{
    if (!bbox_intersect(x1, y1, x2, y2, x3, y3, x4, y4)) {
        *t = INFINITY;
        *u = INFINITY;
        return;
    }
    double tdet = (x1 - x3) * (y3 - y4) - (y1 - y3) * (x3 - x4);
    double udet = (x1 - x2) * (y1 - y3) - (y1 - y2) * (x1 - x3);
    double det = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
    *t = tdet / det;
    *u = -udet / det;
}

/* geometry.py:448-475  two_line_segments_intersect: half-open [0, 1) */
static int seg_hit(double t, double u)
{
    if (0 <= t && t < 1) {
        if (0 <= u && u < 1)
            return 1;
    }
    return 0;
}

/* geometry.py:64-105  point_is_inside_box_2d (inclusive) */
static int point_in_box(double x, double y, double x1, double y1, double x2,
                        double y2)
{
    if (!(x1 <= x && x <= x2))
        return 0;
    if (!(y1 <= y && y <= y2))
        return 0;
    return 1;
}

/* geometry.py:737-829  point_is_inside_polygon (extended winding number).
 * The winding number is a sum of +-1 and +-0.5, exact in any order. */
static int point_in_polygon(double x, double y, const double *vx,
                            const double *vy, int64_t n)
{
    double w = 0;
    for (int64_t v = 0; v < n; v++) {
        int64_t i = v - 1; /* python index -1 wraps to the last vertex */
        int64_t ia = i < 0 ? i + n : i;
        double x0 = vx[ia] - x;
        double y0 = vy[ia] - y;
        double x1 = vx[v] - x;
        double y1 = vy[v] - y;
        if (y0 * y1 < 0) {
            double r = x0 + y0 * (x1 - x0) / (y0 - y1);
            if (r > 0) {
                if (y0 < 0) w += 1; else w -= 1;
            } else {
                if (y0 < 0) w -= 1; else w += 1;
            }
        } else if (y0 == 0) {
            if (x0 > 0) {
                if (y1 > 0) w += 0.5;
                else if (y1 < 0) w -= 0.5;
            } else if (x0 < 0) {
                if (y1 < 0) w += 0.5;
                else if (y1 > 0) w -= 0.5;
            }
        } else if (y1 == 0) {
            if (x1 > 0) {
                if (y0 < 0) w += 0.5;
                else if (y0 > 0) w -= 0.5;
            } else if (x1 < 0) {
                if (y0 > 0) w += 0.5;
                else if (y0 < 0) w -= 0.5;
            }
        }
    }
    return w != 0;
}

/* geometry.py:968-995  area_triangle (signed, origin as third vertex) */
static double area_tri(double x1, double y1, double x2, double y2)
{
    return (x1 * y2 - x2 * y1) / 2;
}

/* exported wrappers for the known-answer tests */
int rgo_bbox_intersect(double xp1, double yp1, double xp2, double yp2,
                       double xq1, double yq1, double xq2, double yq2)
{
    return bbox_intersect(xp1, yp1, xp2, yp2, xq1, yq1, xq2, yq2);
}

void rgo_seg_params(const double *pq, double *tu)
{
    seg_params(pq[0], pq[1], pq[2], pq[3], pq[4], pq[5], pq[6], pq[7], &tu[0],
               &tu[1]);
}

int rgo_seg_hit(double t, double u) { return seg_hit(t, u); }

int rgo_point_in_polygon(double x, double y, const double *vx,
                         const double *vy, int64_t n)
{
    return point_in_polygon(x, y, vx, vy, n);
}

/* ------------------------------------------------------------------------ */
/* _grids.py                                                                */
/* ------------------------------------------------------------------------ */

/* A 2-D vertex array viewed through _arrays.align_axis_right
 * (_arrays.py:128-149): axis 0 transposes, axis 1 is the identity. */
typedef struct {
    const double *p;
    int64_t s0, s1; /* strides of the aligned view (elements) */
    int64_t n0, n1; /* shape of the aligned view */
} view2;

static view2 aligned(const double *p, int64_t nx, int64_t ny, int axis)
{
    view2 v;
    v.p = p;
    if (axis == 0) {
        v.s0 = 1; v.s1 = ny; v.n0 = ny; v.n1 = nx;
    } else {
        v.s0 = ny; v.s1 = 1; v.n0 = nx; v.n1 = ny;
    }
    return v;
}

static inline double at(view2 v, int64_t i, int64_t j)
{
    if (i < 0) i += v.n0;
    if (j < 0) j += v.n1;
    return v.p[i * v.s0 + j * v.s1];
}

/* _grids.py:84-140  _grid_volume_sweep, in the reference's update order.
 * out is (nx-1, ny-1) row-major; its aligned view uses the same rule. */
static void grid_volume_sweep(const double *x, const double *y, int64_t nx,
                              int64_t ny, double *out, int axis)
{
    view2 xa = aligned(x, nx, ny, axis);
    view2 ya = aligned(y, nx, ny, axis);
    if (axis == 0) { /* _grids.py:107-108 swaps x and y for axis 0 */
        view2 s = xa; xa = ya; ya = s;
    }
    int64_t ocx = nx - 1, ocy = ny - 1;
    int64_t os0 = axis == 0 ? 1 : ocy, os1 = axis == 0 ? ocy : 1;
    int64_t num_i = xa.n0, num_j = xa.n1;
    for (int64_t j = 0; j < num_j - 1; j++) {
        for (int64_t i = 0; i < num_i; i++) {
            double x1 = at(xa, i, j), y1 = at(ya, i, j);
            double x2 = at(xa, i, j + 1), y2 = at(ya, i, j + 1);
            double area = area_tri(x1, y1, x2, y2);
            if (i - 1 >= 0)
                out[(i - 1) * os0 + j * os1] += area;
            if (i < num_i - 1)
                out[i * os0 + j * os1] -= area;
        }
    }
    (void)ocx;
}

/* _grids.py:50-75  grid_volume: zeros, then axis 0, then axis 1 */
void rgo_grid_volume(const double *x, const double *y, int64_t nx, int64_t ny,
                     double *out)
{
    memset(out, 0, sizeof(double) * (size_t)((nx - 1) * (ny - 1)));
    grid_volume_sweep(x, y, nx, ny, out, 0);
    grid_volume_sweep(x, y, nx, ny, out, 1);
}

/* _grids.py:167-215  grid_boundary: j=0 (i up), i=last (j up),
 * j=last (i down), i=0 (j down).  bx/by hold 2(nx-1)+2(ny-1) entries. */
int64_t rgo_grid_boundary(const double *x, const double *y, int64_t nx,
                          int64_t ny, double *bx, double *by)
{
    int64_t n = 0;
    for (int64_t i = 0; i < nx - 1; i++) {
        bx[n] = x[i * ny + 0]; by[n] = y[i * ny + 0]; n++;
    }
    for (int64_t j = 0; j < ny - 1; j++) {
        bx[n] = x[(nx - 1) * ny + j]; by[n] = y[(nx - 1) * ny + j]; n++;
    }
    for (int64_t i = 0; i < nx - 1; i++) {
        int64_t ii = nx - 1 - i; /* ~i */
        bx[n] = x[ii * ny + ny - 1]; by[n] = y[ii * ny + ny - 1]; n++;
    }
    for (int64_t j = 0; j < ny - 1; j++) {
        int64_t jj = ny - 1 - j; /* ~j */
        bx[n] = x[0 * ny + jj]; by[n] = y[0 * ny + jj]; n++;
    }
    return n;
}

static int cell_contains(const double *x, const double *y, int64_t ny,
                         int64_t i, int64_t j, double px, double py)
{
    double vx[4], vy[4];
    vx[0] = x[i * ny + j];             vy[0] = y[i * ny + j];
    vx[1] = x[(i + 1) * ny + j];       vy[1] = y[(i + 1) * ny + j];
    vx[2] = x[(i + 1) * ny + j + 1];   vy[2] = y[(i + 1) * ny + j + 1];
    vx[3] = x[i * ny + j + 1];         vy[3] = y[i * ny + j + 1];
    return point_in_polygon(px, py, vx, vy, 4);
}

/* _grids.py:223-279  index_of_point_brute: first cell in (i, j) order */
void rgo_index_of_point_brute(double px, double py, const double *x,
                              const double *y, int64_t nx, int64_t ny,
                              int64_t *ij)
{
    for (int64_t i = 0; i < nx - 1; i++)
        for (int64_t j = 0; j < ny - 1; j++)
            if (cell_contains(x, y, ny, i, j, px, py)) {
                ij[0] = i; ij[1] = j;
                return;
            }
    ij[0] = RGO_MAXSIZE;
    ij[1] = RGO_MAXSIZE;
}

/* _grids.py:286-349  _index_of_point_local: lowest-index cell of the 3x3
 * block around (i0, j0) that contains the point */
static void index_of_point_local(double px, double py, const double *x,
                                 const double *y, int64_t nx, int64_t ny,
                                 int64_t i0, int64_t j0, int64_t *ij)
{
    int64_t ncx = nx - 1, ncy = ny - 1;
    int64_t ia = i0 - 1 > 0 ? i0 - 1 : 0, ib = i0 + 2 < ncx ? i0 + 2 : ncx;
    int64_t ja = j0 - 1 > 0 ? j0 - 1 : 0, jb = j0 + 2 < ncy ? j0 + 2 : ncy;
    for (int64_t ii = ia; ii < ib; ii++)
        for (int64_t jj = ja; jj < jb; jj++)
            if (cell_contains(x, y, ny, ii, jj, px, py)) {
                ij[0] = ii; ij[1] = jj;
                return;
            }
    ij[0] = RGO_MAXSIZE;
    ij[1] = RGO_MAXSIZE;
}

/* floor() converted to int64.  numba's int(math.floor(v)) is undefined for
 * NaN and values beyond int64; those saturate here (parity unpinned there). */
static int64_t floor_i64(double v)
{
    double f = floor(v);
    if (!(f > -4.0e18)) return (int64_t)-4000000000000000000LL;
    if (!(f < 4.0e18)) return (int64_t)4000000000000000000LL;
    return (int64_t)f;
}

/* _interp_ndarray.py:226-269  _bilinear_interpolation (clamped) */
static double bilinear(const double *a, int64_t nx, int64_t ny, double x,
                       double y)
{
    int64_t x00 = floor_i64(x), y00 = floor_i64(y);
    if (x00 < 0) x00 = 0;
    else if (x00 > nx - 2) x00 = nx - 2;
    if (y00 < 0) y00 = 0;
    else if (y00 > ny - 2) y00 = ny - 2;
    double a00 = a[x00 * ny + y00];
    double a01 = a[x00 * ny + y00 + 1];
    double a10 = a[(x00 + 1) * ny + y00];
    double a11 = a[(x00 + 1) * ny + y00 + 1];
    double dx = x - (double)x00;
    double dy = y - (double)y00;
    double w00 = (1 - dx) * (1 - dy);
    double w01 = (1 - dx) * dy;
    double w10 = dx * (1 - dy);
    double w11 = dx * dy;
    return (a00 * w00) + (a01 * w01) + (a10 * w10) + (a11 * w11);
}

/* _grids.py:356-463  index_of_point_secant */
void rgo_index_of_point_secant(double px, double py, const double *x,
                               const double *y, int64_t nx, int64_t ny,
                               int64_t *ij)
{
    const double h = 1e-3;
    int64_t ncx = nx - 1, ncy = ny - 1;
    double i = (double)nx / 2;
    double j = (double)ny / 2;
    for (int it = 0; it < 100; it++) {
        double xv = bilinear(x, nx, ny, i, j);
        double yv = bilinear(y, nx, ny, i, j);
        int64_t i0 = floor_i64(i), j0 = floor_i64(j);
        double ex = xv - px;
        double ey = yv - py;
        if (i0 >= 0 && j0 >= 0 && i0 < ncx && j0 < ncy) {
            if (cell_contains(x, y, ny, i0, j0, px, py)) {
                index_of_point_local(px, py, x, y, nx, ny, i0, j0, ij);
                return;
            }
        }
        if (fabs(ex) < 1e-10 && fabs(ey) < 1e-10) {
            index_of_point_local(px, py, x, y, nx, ny, i0, j0, ij);
            return;
        }
        double dx_di = (bilinear(x, nx, ny, i + h, j) - xv) / h;
        double dx_dj = (bilinear(x, nx, ny, i, j + h) - xv) / h;
        double dy_di = (bilinear(y, nx, ny, i + h, j) - yv) / h;
        double dy_dj = (bilinear(y, nx, ny, i, j + h) - yv) / h;
        double det = dx_di * dy_dj - dx_dj * dy_di;
        if (det == 0) {
            rgo_index_of_point_brute(px, py, x, y, nx, ny, ij);
            return;
        }
        double di = (+dy_dj * ex - dx_dj * ey) / det;
        double dj = (-dy_di * ex + dx_di * ey) / det;
        i -= di;
        j -= dj;
    }
    rgo_index_of_point_brute(px, py, x, y, nx, ny, ij);
}

/* ------------------------------------------------------------------------ */
/* _weights_conservative_2d.py  -- the sweep                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t n, cap;
    int64_t *in, *out;
    double *w;
} tbuf;

static void tbuf_push(tbuf *b, int64_t in, int64_t out, double w)
{
    if (b->n == b->cap) {
        int64_t cap = b->cap ? 2 * b->cap : 64;
        b->in = (int64_t *)realloc(b->in, sizeof(int64_t) * (size_t)cap);
        b->out = (int64_t *)realloc(b->out, sizeof(int64_t) * (size_t)cap);
        b->w = (double *)realloc(b->w, sizeof(double) * (size_t)cap);
        b->cap = cap;
    }
    b->in[b->n] = in;
    b->out[b->n] = out;
    b->w[b->n] = w;
    b->n++;
}

static void tbuf_free(tbuf *b)
{
    free(b->in); free(b->out); free(b->w);
    memset(b, 0, sizeof(*b));
}

/* _grids.py:143-158 */
static const int64_t cell_normals[4][2] = {{-1, 0}, {0, -1}, {1, 0}, {0, 1}};
static const int64_t cell_vertex[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

typedef struct {
    /* the static grid, original orientation */
    const double *x, *y;
    int64_t nx, ny;
    double bx0, by0, bx1, by1; /* bbox (_weights_conservative_2d.py:179-180) */
    const double *bdx, *bdy;   /* boundary polygon */
    int64_t nb;
    /* the sweep grid, aligned view for this axis */
    view2 xs, ys;
    /* shared */
    const double *vol_in;      /* (nxi-1, nyi-1) */
    const double *wi;          /* optional, same shape */
    int64_t ncx_in, ncy_in, ncx_out, ncy_out;
    int sweep_input, axis;
} sweep_ctx;

/* _weights_conservative_2d.py:876-912  _index_input_output + flattening,
 * then 757-830 _calc_and_save_weights */
static void calc_and_save(const sweep_ctx *c, tbuf *b, double x1, double y1,
                          double x2, double y2, int64_t row, int64_t k,
                          int64_t sx, int64_t sy)
{
    double area = area_tri(x1, y1, x2, y2);
    if (c->axis == 0)
        area = -area;
    for (int side = 0; side < 2; side++) {
        int64_t r = side == 0 ? row - 1 : row;
        if (side == 0 && !(r >= 0))
            continue;
        if (side == 1 && !(r < c->xs.n0 - 1))
            continue;
        int64_t si = r, sj = k;
        if (c->axis == 0) { int64_t s = si; si = sj; sj = s; }
        int64_t ii, ij, oi, oj;
        if (c->sweep_input) { ii = si; ij = sj; oi = sx; oj = sy; }
        else { ii = sx; ij = sy; oi = si; oj = sj; }
        double a_in = c->vol_in[ii * c->ncy_in + ij];
        double w = side == 0 ? area / a_in : -area / a_in;
        if (c->wi)
            w *= c->wi[ii * c->ncy_in + ij];
        tbuf_push(b, ii * c->ncy_in + ij, oi * c->ncy_out + oj, w);
    }
}

typedef struct {
    double x1, y1, x2, y2;
    int64_t k, sx, sy, edge_last;
    int outside;
} walk_state;

/* _weights_conservative_2d.py:401-552  _step_outside_static */
static void step_outside(const sweep_ctx *c, walk_state *s)
{
    int64_t edge_in = s->edge_last;
    int64_t edge_new = RGO_MAXSIZE;
    int outside = 1, found = 0;
    double t_min = INFINITY;
    double px1 = s->x1, py1 = s->y1, px2 = s->x2, py2 = s->y2;
    int64_t ncx = c->nx - 1, ncy = c->ny - 1;
    for (int axis = 0; axis < 2; axis++) {
        view2 xa = aligned(c->x, c->nx, c->ny, axis);
        view2 ya = aligned(c->y, c->nx, c->ny, axis);
        int64_t num_i = xa.n0;
        for (int face = 0; face < 2; face++) {
            int64_t j_edge = face > 0 ? -1 : 0;
            for (int64_t i0 = 0; i0 < num_i - 1; i0++) {
                int64_t i1 = i0 + 1;
                double x3 = at(xa, i0, j_edge), y3 = at(ya, i0, j_edge);
                double x4 = at(xa, i1, j_edge), y4 = at(ya, i1, j_edge);
                double t, u;
                seg_params(px1, py1, px2, py2, x3, y3, x4, y4, &t, &u);
                if (seg_hit(t, u)) {
                    if (t < t_min) {
                        int64_t m = i0;
                        int64_t n = face > 0 ? (axis == 0 ? ncx : ncy) - 1 : 0;
                        int64_t isx, isy;
                        if (axis == 0) { isx = n; isy = m; }
                        else { isx = m; isy = n; }
                        int64_t flat = isx * ncy + isy;
                        if (flat == edge_in)
                            continue;
                        found = 1;
                        t_min = t;
                        s->sx = isx;
                        s->sy = isy;
                        s->x2 = px1 + t * (px2 - px1);
                        s->y2 = py1 + t * (py2 - py1);
                        edge_new = face * 2 + axis;
                        outside = 0;
                    }
                }
            }
        }
    }
    if (!found)
        s->k += 1;
    s->edge_last = edge_new;
    s->outside = outside;
}

/* _weights_conservative_2d.py:561-749  _step_inside_static */
static void step_inside(const sweep_ctx *c, tbuf *b, int64_t row,
                        walk_state *s)
{
    double px1 = s->x1, py1 = s->y1, px2 = s->x2, py2 = s->y2;
    int64_t ncx = c->nx - 1, ncy = c->ny - 1;
    int64_t k_new, sx_new, sy_new, edge_new;
    int hit = 0;
    for (int v = 0; v < 4; v++) {
        if ((int64_t)v == s->edge_last)
            continue;
        int vp = (v + 3) % 4; /* indices_cell_vertex[v - 1] */
        int64_t i3 = cell_vertex[vp][0] + s->sx, j3 = cell_vertex[vp][1] + s->sy;
        int64_t i4 = cell_vertex[v][0] + s->sx, j4 = cell_vertex[v][1] + s->sy;
        double x3 = c->x[i3 * c->ny + j3], y3 = c->y[i3 * c->ny + j3];
        double x4 = c->x[i4 * c->ny + j4], y4 = c->y[i4 * c->ny + j4];
        double t, u;
        seg_params(px1, py1, px2, py2, x3, y3, x4, y4, &t, &u);
        if (seg_hit(t, u)) {
            px2 = px1 + t * (s->x2 - px1);
            py2 = py1 + t * (s->y2 - py1);
            k_new = s->k;
            sx_new = s->sx + cell_normals[v][0];
            sy_new = s->sy + cell_normals[v][1];
            edge_new = (v + 2) % 4;
            hit = 1;
            break;
        }
    }
    if (!hit) {
        k_new = s->k + 1;
        sx_new = s->sx;
        sy_new = s->sy;
        edge_new = RGO_MAXSIZE;
    }
    calc_and_save(c, b, px1, py1, px2, py2, row, s->k, s->sx, s->sy);
    s->outside = 0;
    if (!(sx_new >= 0 && sy_new >= 0 && sx_new < ncx && sy_new < ncy)) {
        edge_new = s->sx * ncy + s->sy;
        sx_new = RGO_MAXSIZE;
        sy_new = RGO_MAXSIZE;
        s->outside = 1;
    }
    s->x2 = px2;
    s->y2 = py2;
    s->k = k_new;
    s->sx = sx_new;
    s->sy = sy_new;
    s->edge_last = edge_new;
}

/* _weights_conservative_2d.py:286-387  one row of _sweep_along_axis */
static void sweep_row(const sweep_ctx *c, int64_t row, tbuf *b)
{
    walk_state s;
    s.k = 0;
    s.sx = RGO_MAXSIZE;
    s.sy = RGO_MAXSIZE;
    s.outside = 1;
    s.edge_last = RGO_MAXSIZE;
    s.x1 = at(c->xs, row, 0);
    s.y1 = at(c->ys, row, 0);
    if (point_in_box(s.x1, s.y1, c->bx0, c->by0, c->bx1, c->by1)) {
        if (point_in_polygon(s.x1, s.y1, c->bdx, c->bdy, c->nb)) {
            int64_t ij[2];
            rgo_index_of_point_secant(s.x1, s.y1, c->x, c->y, c->nx, c->ny, ij);
            s.sx = ij[0];
            s.sy = ij[1];
            s.outside = 0;
        }
    }
    int64_t nk = c->xs.n1;
    while (s.k < nk - 1) {
        s.x2 = at(c->xs, row, s.k + 1);
        s.y2 = at(c->ys, row, s.k + 1);
        if (s.outside)
            step_outside(c, &s);
        else
            step_inside(c, b, row, &s);
        s.x1 = s.x2;
        s.y1 = s.y2;
    }
}

static int g_threads = 1;

void rgo_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* _weights_conservative_2d.py:133-203 (_sweep_grid) and 211-392
 * (_sweep_along_axis) for one pass; appends rows in order to *all. */
static void sweep_grid(const double *xi, const double *yi, int64_t nxi,
                       int64_t nyi, const double *xo, const double *yo,
                       int64_t nxo, int64_t nyo, const double *vol_in,
                       const double *wi, int sweep_input, tbuf *all)
{
    sweep_ctx c;
    const double *xsw, *ysw;
    int64_t nxs, nys;
    if (sweep_input) {
        xsw = xi; ysw = yi; nxs = nxi; nys = nyi;
        c.x = xo; c.y = yo; c.nx = nxo; c.ny = nyo;
    } else {
        xsw = xo; ysw = yo; nxs = nxo; nys = nyo;
        c.x = xi; c.y = yi; c.nx = nxi; c.ny = nyi;
    }
    int64_t nstat = c.nx * c.ny;
    c.bx0 = c.bx1 = c.x[0];
    c.by0 = c.by1 = c.y[0];
    for (int64_t q = 1; q < nstat; q++) {
        /* numpy min/max: NaN propagates; the grids here are finite */
        if (c.x[q] < c.bx0) c.bx0 = c.x[q];
        if (c.x[q] > c.bx1) c.bx1 = c.x[q];
        if (c.y[q] < c.by0) c.by0 = c.y[q];
        if (c.y[q] > c.by1) c.by1 = c.y[q];
    }
    int64_t nb = 2 * (c.nx - 1) + 2 * (c.ny - 1);
    double *bdx = (double *)malloc(sizeof(double) * (size_t)nb);
    double *bdy = (double *)malloc(sizeof(double) * (size_t)nb);
    rgo_grid_boundary(c.x, c.y, c.nx, c.ny, bdx, bdy);
    c.bdx = bdx; c.bdy = bdy; c.nb = nb;
    c.vol_in = vol_in;
    c.wi = wi;
    c.ncx_in = nxi - 1; c.ncy_in = nyi - 1;
    c.ncx_out = nxo - 1; c.ncy_out = nyo - 1;
    c.sweep_input = sweep_input;
    for (int axis = 0; axis < 2; axis++) {
        c.axis = axis;
        c.xs = aligned(xsw, nxs, nys, axis);
        c.ys = aligned(ysw, nxs, nys, axis);
        int64_t nrows = c.xs.n0;
        tbuf *rows = (tbuf *)calloc((size_t)nrows, sizeof(tbuf));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
#endif
        for (int64_t r = 0; r < nrows; r++)
            sweep_row(&c, r, &rows[r]);
        for (int64_t r = 0; r < nrows; r++) {
            for (int64_t q = 0; q < rows[r].n; q++)
                tbuf_push(all, rows[r].in[q], rows[r].out[q], rows[r].w[q]);
            tbuf_free(&rows[r]);
        }
        free(rows);
    }
    free(bdx);
    free(bdy);
}

/* _weights_conservative_2d.py:80-126  weights_conservative_2d: uncoalesced
 * triples in emission order.  Returns the count; the three arrays are malloc-ed
 * and released with rgo_free. */
int64_t rgo_sweep_2d(const double *xi, const double *yi, int64_t nxi,
                     int64_t nyi, const double *xo, const double *yo,
                     int64_t nxo, int64_t nyo, const double *wi,
                     int64_t **in, int64_t **out, double **w)
{
    double *vol = (double *)malloc(sizeof(double) *
                                   (size_t)((nxi - 1) * (nyi - 1)));
    rgo_grid_volume(xi, yi, nxi, nyi, vol);
    tbuf all;
    memset(&all, 0, sizeof(all));
    sweep_grid(xi, yi, nxi, nyi, xo, yo, nxo, nyo, vol, wi, 0, &all);
    sweep_grid(xi, yi, nxi, nyi, xo, yo, nxo, nyo, vol, wi, 1, &all);
    free(vol);
    *in = all.in;
    *out = all.out;
    *w = all.w;
    return all.n;
}

void rgo_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* _weights_arrays.py:44-73  _coalesce                                      */
/* ------------------------------------------------------------------------ */

/* numpy's pairwise summation (the reduction inner loop that
 * np.add.reduceat uses for float64): below 8 terms a sequential sum seeded
 * with -0.0; up to 128 terms eight interleaved accumulators; above that a
 * split at n/2 rounded down to a multiple of 8. */
static double pairwise_sum(const double *a, int64_t n)
{
    if (n < 8) {
        double res = -0.0;
        for (int64_t i = 0; i < n; i++)
            res += a[i];
        return res;
    } else if (n <= 128) {
        double r[8];
        for (int q = 0; q < 8; q++)
            r[q] = a[q];
        int64_t i;
        for (i = 8; i < n - (n % 8); i += 8)
            for (int q = 0; q < 8; q++)
                r[q] += a[i + q];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) +
                     ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; i++)
            res += a[i];
        return res;
    } else {
        int64_t n2 = n / 2;
        n2 -= n2 % 8;
        return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
    }
}

/* np.add.reduceat on one segment: the first element seeds the output, the
 * rest is reduced pairwise and added to it. */
double rgo_reduce_segment(const double *a, int64_t n)
{
    if (n == 1)
        return a[0];
    return a[0] + pairwise_sum(a + 1, n - 1);
}

typedef struct {
    int64_t key, pos;
} keypos;

static int keypos_cmp(const void *pa, const void *pb)
{
    const keypos *a = (const keypos *)pa, *b = (const keypos *)pb;
    if (a->key != b->key)
        return a->key < b->key ? -1 : 1;
    return a->pos < b->pos ? -1 : (a->pos > b->pos ? 1 : 0);
}

/* Coalesce: key = (in - min_in) * span_out + (out - min_out), stable argsort,
 * runs of equal keys reduced with numpy's reduceat order.  Output arrays
 * (capacity n) are ascending by key.  Returns the unique count. */
int64_t rgo_coalesce(int64_t n, const int64_t *in, const int64_t *out,
                     const double *w, int64_t *in_c, int64_t *out_c,
                     double *w_c)
{
    if (n == 0)
        return 0;
    int64_t min_in = in[0], min_out = out[0], max_out = out[0];
    for (int64_t q = 1; q < n; q++) {
        if (in[q] < min_in) min_in = in[q];
        if (out[q] < min_out) min_out = out[q];
        if (out[q] > max_out) max_out = out[q];
    }
    int64_t span = max_out - min_out + 1;
    keypos *kp = (keypos *)malloc(sizeof(keypos) * (size_t)n);
    for (int64_t q = 0; q < n; q++) {
        kp[q].key = (in[q] - min_in) * span + (out[q] - min_out);
        kp[q].pos = q;
    }
    /* (key, original position) is a total order: same result as a stable
     * sort by key */
    qsort(kp, (size_t)n, sizeof(keypos), keypos_cmp);
    double *vals = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t q = 0; q < n; q++)
        vals[q] = w[kp[q].pos];
    int64_t m = 0, s = 0;
    while (s < n) {
        int64_t e = s + 1;
        while (e < n && kp[e].key == kp[s].key)
            e++;
        in_c[m] = kp[s].key / span + min_in;
        out_c[m] = kp[s].key % span + min_out;
        w_c[m] = rgo_reduce_segment(vals + s, e - s);
        m++;
        s = e;
    }
    free(vals);
    free(kp);
    return m;
}

/* ------------------------------------------------------------------------ */
/* _regrid_from_weights.py:165-182  _regrid_from_weights                    */
/* ------------------------------------------------------------------------ */

/* values_out (D, n_out) must be zeroed by the caller, as rfw:111-118 does.
 * Elements d run in parallel (prange over d); each is a sequential
 * scatter in storage order with a separate multiply and add. */
void rgo_apply(int64_t nnz, const int64_t *in, const int64_t *out,
               const double *w, int64_t D, const double *vin, int64_t n_in,
               double *vout, int64_t n_out)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_threads)
#endif
    for (int64_t d = 0; d < D; d++) {
        const double *x = vin + d * n_in;
        double *y = vout + d * n_out;
        for (int64_t q = 0; q < nnz; q++) {
            /* numba indexing wraps negative indices (1-D builders emit them
             * for descending grids) */
            const int64_t a = in[q] < 0 ? in[q] + n_in : in[q];
            const int64_t b = out[q] < 0 ? out[q] + n_out : out[q];
            double p = w[q] * x[a];
            y[b] += p;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* _find_indices: 1-D brute force and searchsorted                          */
/* ------------------------------------------------------------------------ */

/* _find_indices_brute.py:24-51 */
void rgo_find_indices_brute_1d(const double *xin, int64_t m,
                               const double *xout, int64_t n, int64_t D,
                               int64_t fill, int64_t *res)
{
    for (int64_t d = 0; d < D; d++) {
        const double *a = xin + d * m;
        for (int64_t i = 0; i < n; i++) {
            double v = xout[d * n + i];
            int64_t r = fill;
            for (int64_t q = 0; q < m - 1; q++)
                if (a[q] <= v && v <= a[q + 1]) {
                    r = q;
                    break;
                }
            res[d * n + i] = r;
        }
    }
}

/* numpy's ordering for searchsorted: NaN sorts after every number */
static int np_less(double a, double b)
{
    return a < b || (b != b && a == a);
}

/* _find_indices_searchsorted.py:24-62, with np.searchsorted(side="left") */
void rgo_find_indices_searchsorted_1d(const double *xin, int64_t m,
                                      const double *xout, int64_t n,
                                      int64_t D, int64_t fill, int64_t *res)
{
    for (int64_t d = 0; d < D; d++) {
        const double *a = xin + d * m;
        double amin = a[0];
        for (int64_t i = 0; i < n; i++) {
            double v = xout[d * n + i];
            int64_t lo = 0, hi = m;
            while (lo < hi) {
                int64_t mid = lo + (hi - lo) / 2;
                if (np_less(a[mid], v)) lo = mid + 1;
                else hi = mid;
            }
            int64_t r = lo - 1;
            if (v == amin) r = 0;
            else if (r < 0) r = fill;
            else if (r > m - 2) r = fill;
            res[d * n + i] = r;
        }
    }
}

/* _fill/_gauss_seidel.py:83-139: red-black Gauss-Seidel, each colour swept
 * row-major with periodic neighbours, T independent (ny, nx) slices */
void rgo_fill_gauss_seidel_2d(double *a, const uint8_t *where, int64_t T,
                              int64_t ny, int64_t nx, int64_t num_iterations)
{
    const double xmin = -1, xmax = 1, ymin = -1, ymax = 1;
    const double dx = (xmax - xmin) / (double)(nx - 1);
    const double dy = (ymax - ymin) / (double)(ny - 1);
    const double dxxinv = 1 / (dx * dx), dyyinv = 1 / (dy * dy);
    const double dcent = 1 / (2 * (dxxinv + dyyinv));
    for (int64_t t = 0; t < T; t++) {
        double *at = a + t * ny * nx;
        const uint8_t *wt = where + t * ny * nx;
        for (int64_t k = 0; k < num_iterations; k++) {
            for (int is_odd = 0; is_odd < 2; is_odd++) {
                for (int64_t j = 0; j < ny; j++) {
                    for (int64_t i = 0; i < nx; i++) {
                        if (((i + j) & 1) != is_odd || !wt[j * nx + i]) continue;
                        int64_t i9 = ((i - 1) % nx + nx) % nx, i1 = (i + 1) % nx;
                        int64_t j9 = ((j - 1) % ny + ny) % ny, j1 = (j + 1) % ny;
                        double xterm = dxxinv * (at[j * nx + i9] + at[j * nx + i1]);
                        double yterm = dyyinv * (at[j9 * nx + i] + at[j1 * nx + i]);
                        at[j * nx + i] = (xterm + yterm) * dcent;
                    }
                }
            }
        }
    }
}

/* _interp_ndarray.py:208-224 (_linear_interpolation), D rows of n samples,
 * m points per row */
void rgo_interp_linear_1d(const double *a, int64_t n, const double *x, int64_t m,
                          int64_t D, double *out)
{
    for (int64_t d = 0; d < D; d++) {
        for (int64_t q = 0; q < m; q++) {
            double xv = x[d * m + q];
            int64_t x0 = (int64_t)floor(xv);
            if (x0 < 0) x0 = 0;
            else if (x0 > n - 2) x0 = n - 2;
            double a0 = a[d * n + x0], a1 = a[d * n + x0 + 1];
            double dxv = xv - (double)x0;
            out[d * m + q] = a0 * (1 - dxv) + a1 * dxv;
        }
    }
}

/* _interp_ndarray.py:227-269 (_bilinear_interpolation) in the written order
 * of its final sum, D arrays of (nx, ny), m points per array */
void rgo_interp_linear_2d(const double *a, int64_t nx, int64_t ny, const double *x,
                          const double *y, int64_t m, int64_t D, double *out)
{
    for (int64_t d = 0; d < D; d++) {
        const double *ad = a + d * nx * ny;
        for (int64_t q = 0; q < m; q++) {
            double xv = x[d * m + q], yv = y[d * m + q];
            int64_t x00 = (int64_t)floor(xv), y00 = (int64_t)floor(yv);
            if (x00 < 0) x00 = 0;
            else if (x00 > nx - 2) x00 = nx - 2;
            if (y00 < 0) y00 = 0;
            else if (y00 > ny - 2) y00 = ny - 2;
            double a00 = ad[x00 * ny + y00], a01 = ad[x00 * ny + y00 + 1];
            double a10 = ad[(x00 + 1) * ny + y00], a11 = ad[(x00 + 1) * ny + y00 + 1];
            double dxv = xv - (double)x00, dyv = yv - (double)y00;
            double w00 = (1 - dxv) * (1 - dyv), w01 = (1 - dxv) * dyv;
            double w10 = dxv * (1 - dyv), w11 = dxv * dyv;
            out[d * m + q] = (a00 * w00) + (a01 * w01) + (a10 * w10) + (a11 * w11);
        }
    }
}
