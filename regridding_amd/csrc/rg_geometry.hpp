// rg_geometry.hpp -- FP64 geometry primitives of the conservative sweep.
//
// Every function here is __host__ __device__ so that the GPU kernels and the
// CPU emulation harness (tests/emu) execute the same instructions.  The build
// uses -ffp-contract=off and no fast-math: each product and each quotient is
// rounded separately, in the order the reference writes the expression, so
// the discrete decisions of the sweep (hit / no hit, t ordering) and the
// emitted weights are bit-identical to an exact-order run of the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RG_HD __host__ __device__ __forceinline__

namespace rg {

// Element i of an array in device global memory.  A pointer read from a
// table in device memory (the grids of a Build) carries no address space,
// so plain loads through it are flat loads, whose completion is also counted
// on the LDS / scalar-load counter: every wait for one then waits for both
// kinds.  On the device the load names the global address space; the host
// twin serves the CPU emulation of the same functions.
template <class T, class I>
__device__ __forceinline__ T gload(const T* p, I i)
{
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <class T, class I>
__host__ __forceinline__ T gload(const T* p, I i)
{
    return p[i];
}

// geometry.py:153-284 bounding_boxes_intersect_2d: inclusive (touching
// boxes intersect); each box is given by two corner points in any order.
RG_HD bool bbox_overlap(double xp1, double yp1, double xp2, double yp2,
                        double xq1, double yq1, double xq2, double yq2)
{
    double pxl = xp1 > xp2 ? xp2 : xp1, pxh = xp1 > xp2 ? xp1 : xp2;
    double qxl = xq1 > xq2 ? xq2 : xq1, qxh = xq1 > xq2 ? xq1 : xq2;
    if (!(pxl <= qxh && qxl <= pxh)) return false;
    double pyl = yp1 > yp2 ? yp2 : yp1, pyh = yp1 > yp2 ? yp1 : yp2;
    double qyl = yq1 > yq2 ? yq2 : yq1, qyh = yq1 > yq2 ? yq1 : yq2;
    if (!(pyl <= qyh && qyl <= pyh)) return false;
    return true;
}

// Intersection parameters of segments p1->p2 and q1->q2
// (geometry.py:370-445): t along p, u along q.  Disjoint bounding boxes give
// (inf, inf); parallel segments give inf/NaN through the division.
struct TU {
    double t, u;
};

RG_HD TU seg_params(double x1, double y1, double x2, double y2, double x3,
                    double y3, double x4, double y4)
{
    TU r;
    if (!bbox_overlap(x1, y1, x2, y2, x3, y3, x4, y4)) {
        r.t = __builtin_huge_val();
        r.u = __builtin_huge_val();
        return r;
    }
    double tdet = (x1 - x3) * (y3 - y4) - (y1 - y3) * (x3 - x4);
    double udet = (x1 - x2) * (y1 - y3) - (y1 - y2) * (x1 - x3);
    double det = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
    r.t = tdet / det;
    r.u = -udet / det;
    return r;
}

// geometry.py:448-475: hit iff 0 <= t < 1 and 0 <= u < 1.
RG_HD bool seg_hit(TU p)
{
    return (0 <= p.t && p.t < 1) && (0 <= p.u && p.u < 1);
}

// geometry.py:968-995 signed area of the triangle (origin, v1, v2).
RG_HD double area_tri(double x1, double y1, double x2, double y2)
{
    return (x1 * y2 - x2 * y1) / 2;
}

// Contribution of polygon edge (vx0,vy0)->(vx1,vy1) to the extended winding
// number of point (x, y) (geometry.py:775-829).  The total is a sum of
// +-1 and +-0.5 terms, so it is exact in any summation order, which lets a
// wavefront split the edges between its lanes.
RG_HD double winding_term(double x, double y, double vx0, double vy0,
                          double vx1, double vy1)
{
    double x0 = vx0 - x, y0 = vy0 - y;
    double x1 = vx1 - x, y1 = vy1 - y;
    if (y0 * y1 < 0) {
        double r = x0 + y0 * (x1 - x0) / (y0 - y1);
        if (r > 0) return y0 < 0 ? 1.0 : -1.0;
        return y0 < 0 ? -1.0 : 1.0;
    }
    if (y0 == 0) {
        if (x0 > 0) {
            if (y1 > 0) return 0.5;
            if (y1 < 0) return -0.5;
        } else if (x0 < 0) {
            if (y1 < 0) return 0.5;
            if (y1 > 0) return -0.5;
        }
        return 0.0;
    }
    if (y1 == 0) {
        if (x1 > 0) {
            if (y0 < 0) return 0.5;
            if (y0 > 0) return -0.5;
        } else if (x1 < 0) {
            if (y0 > 0) return 0.5;
            if (y0 < 0) return -0.5;
        }
    }
    return 0.0;
}

// point_is_inside_polygon for the quadrilateral cell (i, j) of a vertex grid
// with ny vertices per row; vertex order (i,j) (i+1,j) (i+1,j+1) (i,j+1) as in
// _grids.py:256-269.  The polygon loop starts with the closing edge
// v3 -> v0 (python index -1), exactly like geometry.py:775.
RG_HD bool cell_contains(const double* __restrict__ x,
                         const double* __restrict__ y, int64_t ny, int64_t i,
                         int64_t j, double px, double py)
{
    int64_t a = i * ny + j, b = (i + 1) * ny + j;
    double vx0 = x[a], vy0 = y[a];
    double vx1 = x[b], vy1 = y[b];
    double vx2 = x[b + 1], vy2 = y[b + 1];
    double vx3 = x[a + 1], vy3 = y[a + 1];
    double w = 0;
    w += winding_term(px, py, vx3, vy3, vx0, vy0);
    w += winding_term(px, py, vx0, vy0, vx1, vy1);
    w += winding_term(px, py, vx1, vy1, vx2, vy2);
    w += winding_term(px, py, vx2, vy2, vx3, vy3);
    return w != 0;
}

// floor(v) as int64, saturating where numba's int(math.floor(v)) would be
// undefined (NaN, beyond int64).
RG_HD int64_t floor_i64(double v)
{
    double f = __builtin_floor(v);
    if (!(f > -4.0e18)) return -4000000000000000000LL;
    if (!(f < 4.0e18)) return 4000000000000000000LL;
    return (int64_t)f;
}

// _interp_ndarray.py:226-269 _bilinear_interpolation: the lower-left corner
// is clamped into the grid, the fractional offsets are not.
RG_HD double bilinear(const double* __restrict__ a, int64_t nx, int64_t ny,
                      double x, double y)
{
    int64_t x0 = floor_i64(x), y0 = floor_i64(y);
    if (x0 < 0) x0 = 0;
    else if (x0 > nx - 2) x0 = nx - 2;
    if (y0 < 0) y0 = 0;
    else if (y0 > ny - 2) y0 = ny - 2;
    double a00 = a[x0 * ny + y0];
    double a01 = a[x0 * ny + y0 + 1];
    double a10 = a[(x0 + 1) * ny + y0];
    double a11 = a[(x0 + 1) * ny + y0 + 1];
    double dx = x - (double)x0;
    double dy = y - (double)y0;
    double w00 = (1 - dx) * (1 - dy);
    double w01 = (1 - dx) * dy;
    double w10 = dx * (1 - dy);
    double w11 = dx * dy;
    return (a00 * w00) + (a01 * w01) + (a10 * w10) + (a11 * w11);
}

// _grids.py:286-349 _index_of_point_local: lowest-index cell of the 3x3
// block around (i0, j0) containing the point; -1 when none does.
RG_HD int64_t index_local(const double* __restrict__ x,
                          const double* __restrict__ y, int64_t nx, int64_t ny,
                          int64_t i0, int64_t j0, double px, double py)
{
    int64_t ncx = nx - 1, ncy = ny - 1;
    int64_t ia = i0 - 1 > 0 ? i0 - 1 : 0, ib = i0 + 2 < ncx ? i0 + 2 : ncx;
    int64_t ja = j0 - 1 > 0 ? j0 - 1 : 0, jb = j0 + 2 < ncy ? j0 + 2 : ncy;
    for (int64_t ii = ia; ii < ib; ii++)
        for (int64_t jj = ja; jj < jb; jj++)
            if (cell_contains(x, y, ny, ii, jj, px, py)) return ii * ncy + jj;
    return -1;
}

// Result codes of the secant search.
enum : int64_t {
    RG_SEARCH_OUTSIDE = -1,  // sys.maxsize in the reference
    RG_SEARCH_BRUTE = -2,    // the reference falls back to index_of_point_brute
};

// _grids.py:356-463 index_of_point_secant without its brute-force fallback:
// returns the flat cell index, RG_SEARCH_OUTSIDE, or RG_SEARCH_BRUTE when
// the reference would scan every cell (singular Jacobian or 100 iterations
// without a decision).  The caller runs the exhaustive search cooperatively.
RG_HD int64_t index_secant(const double* __restrict__ x,
                           const double* __restrict__ y, int64_t nx, int64_t ny,
                           double px, double py)
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
        if (i0 >= 0 && j0 >= 0 && i0 < ncx && j0 < ncy &&
            cell_contains(x, y, ny, i0, j0, px, py)) {
            int64_t c = index_local(x, y, nx, ny, i0, j0, px, py);
            return c >= 0 ? c : RG_SEARCH_OUTSIDE;
        }
        if (__builtin_fabs(ex) < 1e-10 && __builtin_fabs(ey) < 1e-10) {
            int64_t c = index_local(x, y, nx, ny, i0, j0, px, py);
            return c >= 0 ? c : RG_SEARCH_OUTSIDE;
        }
        double dx_di = (bilinear(x, nx, ny, i + h, j) - xv) / h;
        double dx_dj = (bilinear(x, nx, ny, i, j + h) - xv) / h;
        double dy_di = (bilinear(y, nx, ny, i + h, j) - yv) / h;
        double dy_dj = (bilinear(y, nx, ny, i, j + h) - yv) / h;
        double det = dx_di * dy_dj - dx_dj * dy_di;
        if (det == 0) return RG_SEARCH_BRUTE;
        double di = (+dy_dj * ex - dx_dj * ey) / det;
        double dj = (-dy_di * ex + dx_di * ey) / det;
        i -= di;
        j -= dj;
    }
    return RG_SEARCH_BRUTE;
}

// The cell search over an interleaved (x, y) vertex array: the same
// arithmetic as bilinear / cell_contains / index_local / index_secant on
// the separate arrays, with one 16-byte load per vertex instead of two.
RG_HD double2 bilinear_xy(const double2* __restrict__ a, int64_t nx, int64_t ny, double x,
                          double y)
{
    int64_t x0 = floor_i64(x), y0 = floor_i64(y);
    if (x0 < 0) x0 = 0;
    else if (x0 > nx - 2) x0 = nx - 2;
    if (y0 < 0) y0 = 0;
    else if (y0 > ny - 2) y0 = ny - 2;
    const double2 a00 = gload(a, x0 * ny + y0);
    const double2 a01 = gload(a, x0 * ny + y0 + 1);
    const double2 a10 = gload(a, (x0 + 1) * ny + y0);
    const double2 a11 = gload(a, (x0 + 1) * ny + y0 + 1);
    double dx = x - (double)x0;
    double dy = y - (double)y0;
    double w00 = (1 - dx) * (1 - dy);
    double w01 = (1 - dx) * dy;
    double w10 = dx * (1 - dy);
    double w11 = dx * dy;
    double2 r;
    r.x = (a00.x * w00) + (a01.x * w01) + (a10.x * w10) + (a11.x * w11);
    r.y = (a00.y * w00) + (a01.y * w01) + (a10.y * w10) + (a11.y * w11);
    return r;
}

RG_HD bool cell_contains_xy(const double2* __restrict__ xy, int64_t ny, int64_t i, int64_t j,
                            double px, double py)
{
    int64_t a = i * ny + j, b = (i + 1) * ny + j;
    const double2 v0 = gload(xy, a), v1 = gload(xy, b), v2 = gload(xy, b + 1),
                  v3 = gload(xy, a + 1);
    double w = 0;
    w += winding_term(px, py, v3.x, v3.y, v0.x, v0.y);
    w += winding_term(px, py, v0.x, v0.y, v1.x, v1.y);
    w += winding_term(px, py, v1.x, v1.y, v2.x, v2.y);
    w += winding_term(px, py, v2.x, v2.y, v3.x, v3.y);
    return w != 0;
}

RG_HD int64_t index_local_xy(const double2* __restrict__ xy, int64_t nx, int64_t ny,
                             int64_t i0, int64_t j0, double px, double py)
{
    int64_t ncx = nx - 1, ncy = ny - 1;
    int64_t ia = i0 - 1 > 0 ? i0 - 1 : 0, ib = i0 + 2 < ncx ? i0 + 2 : ncx;
    int64_t ja = j0 - 1 > 0 ? j0 - 1 : 0, jb = j0 + 2 < ncy ? j0 + 2 : ncy;
    for (int64_t ii = ia; ii < ib; ii++)
        for (int64_t jj = ja; jj < jb; jj++)
            if (cell_contains_xy(xy, ny, ii, jj, px, py)) return ii * ncy + jj;
    return -1;
}

RG_HD int64_t index_secant_xy(const double2* __restrict__ xy, int64_t nx, int64_t ny,
                              double px, double py)
{
    const double h = 1e-3;
    int64_t ncx = nx - 1, ncy = ny - 1;
    double i = (double)nx / 2;
    double j = (double)ny / 2;
    for (int it = 0; it < 100; it++) {
        const double2 v = bilinear_xy(xy, nx, ny, i, j);
        const double xv = v.x, yv = v.y;
        int64_t i0 = floor_i64(i), j0 = floor_i64(j);
        double ex = xv - px;
        double ey = yv - py;
        if (i0 >= 0 && j0 >= 0 && i0 < ncx && j0 < ncy &&
            cell_contains_xy(xy, ny, i0, j0, px, py)) {
            int64_t c = index_local_xy(xy, nx, ny, i0, j0, px, py);
            return c >= 0 ? c : RG_SEARCH_OUTSIDE;
        }
        if (__builtin_fabs(ex) < 1e-10 && __builtin_fabs(ey) < 1e-10) {
            int64_t c = index_local_xy(xy, nx, ny, i0, j0, px, py);
            return c >= 0 ? c : RG_SEARCH_OUTSIDE;
        }
        const double2 vi = bilinear_xy(xy, nx, ny, i + h, j);
        const double2 vj = bilinear_xy(xy, nx, ny, i, j + h);
        double dx_di = (vi.x - xv) / h;
        double dx_dj = (vj.x - xv) / h;
        double dy_di = (vi.y - yv) / h;
        double dy_dj = (vj.y - yv) / h;
        double det = dx_di * dy_dj - dx_dj * dy_di;
        if (det == 0) return RG_SEARCH_BRUTE;
        double di = (+dy_dj * ex - dx_dj * ey) / det;
        double dj = (-dy_di * ex + dx_di * ey) / det;
        i -= di;
        j -= dj;
    }
    return RG_SEARCH_BRUTE;
}

}  // namespace rg
