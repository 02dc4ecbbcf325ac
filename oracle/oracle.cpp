// oracle.cpp -- CPU restatement of MGmol's orbital-iteration hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Built with -ffp-contract=off so that every double expression rounds exactly
// as written (the reference is compiled without FMA contraction guarantees; we
// pick the unfused order that the source text states).  Each function cites
// the reference file:line whose arithmetic it restates.
//
// Parity pins (tests/test_oracle_*.py): reference unit tests testLaph4,
// testBatchLaph4, testMGkernels, testGramMatrix restated; Fourier symbols of
// every stencil; R = P^T/8 adjointness; Jacobi amplification per Fourier mode;
// constant-input closed form of the whole V-cycle; box family == grid family
// with 0/1 masks; brute-force / numpy for the dense products.
#include "oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace
{

const double inv12 = 1. / 12.;
const double inv64 = 1. / 64.;

struct G
{
    int d[3];
    int ng;
    double h[3];
    long incx, incy, sizeg, size;
    explicit G(const or_grid_t* g)
    {
        for (int i = 0; i < 3; i++)
        {
            d[i] = g->dim[i];
            h[i] = g->h[i];
        }
        ng    = g->ghosts;
        incy  = d[2] + 2 * ng;
        incx  = (long)(d[1] + 2 * ng) * incy;
        sizeg = (long)(d[0] + 2 * ng) * incx;
        size  = (long)d[0] * d[1] * d[2];
    }
    double inv_h2(int i) const { return 1. / (h[i] * h[i]); }
    long idx(int ix, int iy, int iz) const
    {
        return (long)(ix + ng) * incx + (long)(iy + ng) * incy + iz + ng;
    }
};

inline int wrap(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------
// Halo fill.  GridFuncVector::trade_boundaries, single-task branches
// (src/pb/GridFuncVector.cc:1544-1622, 585-604, 1052-1070, 1426-1440):
// along a Dirichlet direction GridFunc::setBoundaryValues
// (src/pb/GridFunc.cc:2189-2340) zeroes the ghosts on both sides AND the first
// interior layer (index 0: the boundary plane of the vertex-centred grid, so
// that the free points 1 .. n-1 are an odd number); periodic directions copy
// Y then Z then X so that edges and corners hold the periodic image.  The net
// value of every point is zero if a Dirichlet coordinate lies outside
// 1 .. n-1, else (ghosts) the periodic image of the interior point.  The
// reference's loops skip the index-0 layers of the other directions when a
// grid mixes periodic and Dirichlet directions (a boundary row is left
// unset); this restatement zeroes the whole layer, as for an all-Dirichlet
// grid, where the two agree.  bc 3 (no counterpart in the reference's
// GridFunc; used for a localized orbital's box cut out of the grid) gives
// zero ghosts and leaves the interior as it is.
template <typename T>
int trade(const or_grid_t* gg, T* u, long nfunc, const int bc[3])
{
    G g(gg);
    for (int i = 0; i < 3; i++)
        if (g.ng > g.d[i]) return -1;
    const int n0 = g.d[0] + 2 * g.ng, n1 = g.d[1] + 2 * g.ng, n2 = g.d[2] + 2 * g.ng;
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int x = 0; x < n0; x++)
        {
            T* uf = u + f * g.sizeg;
            for (int y = 0; y < n1; y++)
                for (int z = 0; z < n2; z++)
                {
                    int c[3] = { x - g.ng, y - g.ng, z - g.ng };
                    bool ghost = false, zero = false;
                    for (int k = 0; k < 3; k++)
                    {
                        if (bc[k] == 0 && c[k] == 0) zero = true;
                        if (c[k] < 0 || c[k] >= g.d[k])
                        {
                            ghost = true;
                            if (bc[k] == 1)
                                c[k] = wrap(c[k], g.d[k]);
                            else
                                zero = true;
                        }
                    }
                    if (!ghost && !zero) continue;
                    long dst = (long)x * g.incx + (long)y * g.incy + z;
                    uf[dst]  = zero ? (T)0 : uf[g.idx(c[0], c[1], c[2])];
                }
        }
    return 0;
}

// ---------------------------------------------------------------------------
// -Laplacian, 2nd order.  FDkernelDel2_2nd, src/pb/FDkernels.cc:19-75.
// Pair sums are formed in T, the rest in double.
template <typename T>
void del2_2nd(const G& g, const T* v, T* b, long nfunc)
{
    const double c1x = -g.inv_h2(0), c1y = -g.inv_h2(1), c1z = -g.inv_h2(2);
    const double c0  = -2. * (c1x + c1y + c1z);
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                {
                    long i = f * g.sizeg + g.idx(ix, iy, iz);
                    b[i] = (T)(c0 * v[i] + c1x * (T)(v[i - g.incx] + v[i + g.incx])
                               + c1y * (T)(v[i - g.incy] + v[i + g.incy])
                               + c1z * (T)(v[i - 1] + v[i + 1]));
                }
}

struct C4
{
    double c0, c1x, c2x, c1y, c2y, c1z, c2z;
    explicit C4(const double h[3])
    {
        double ih[3];
        for (int i = 0; i < 3; i++)
            ih[i] = 1. / (h[i] * h[i]);
        const double cc0 = inv12 * ih[0];
        c1x              = -16. * cc0;
        c2x              = 1. * cc0;
        const double cc1 = inv12 * ih[1];
        c1y              = -16. * cc1;
        c2y              = 1. * cc1;
        const double cc2 = inv12 * ih[2];
        c1z              = -16. * cc2;
        c2z              = 1. * cc2;
        c0               = -2. * (c1x + c2x + c1y + c2y + c1z + c2z);
    }
};

// -Laplacian, 4th order, 13 points.  FDkernelDel2_4th,
// src/pb/FDkernels.cc:78-162 (coefficients :92-104, loop body :140-156).
// With pot != nullptr this is FDoper::del2_4th_withPot,
// src/pb/FDoper.cc:322-399: (c0 + pot) on the diagonal, output packed.
template <typename T>
inline double lap4_point(const C4& c, const T* a, long i, long incx, long incy, double diag)
{
    return diag * (double)a[i] + c.c1x * ((double)a[i - incx] + (double)a[i + incx])
           + c.c1y * ((double)a[i - incy] + (double)a[i + incy])
           + c.c1z * ((double)a[i - 1] + (double)a[i + 1])
           + c.c2x * ((double)a[i - 2 * incx] + (double)a[i + 2 * incx])
           + c.c2y * ((double)a[i - 2 * incy] + (double)a[i + 2 * incy])
           + c.c2z * ((double)a[i - 2] + (double)a[i + 2]);
}

template <typename T>
void del2_4th(const G& g, const T* a, T* b, long nfunc)
{
    C4 c(g.h);
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                {
                    long i = f * g.sizeg + g.idx(ix, iy, iz);
                    b[i]   = (T)lap4_point(c, a, i, g.incx, g.incy, c.c0);
                }
}

template <typename T>
void del2_4th_withpot(const G& g, const T* a, const double* pot, T* b, long ld, long nfunc)
{
    C4 c(g.h);
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
            {
                long ipx = ((long)ix * g.d[1] + iy) * g.d[2];
                for (int iz = 0; iz < g.d[2]; iz++, ipx++)
                {
                    long i         = f * g.sizeg + g.idx(ix, iy, iz);
                    b[f * ld + ipx] = (T)lap4_point(c, a, i, g.incx, g.incy, c.c0 + pot[ipx]);
                }
            }
}

struct CM
{
    double c0, cx, cy, cz, cxy, cyz, cxz;
    explicit CM(const double h[3])
    {
        double ih[3];
        for (int i = 0; i < 3; i++)
            ih[i] = 1. / (h[i] * h[i]);
        c0  = 16. * inv12 * (ih[0] + ih[1] + ih[2]);
        cx  = -10. * inv12 * ih[0] + 0.125 * c0;
        cy  = -10. * inv12 * ih[1] + 0.125 * c0;
        cz  = -10. * inv12 * ih[2] + 0.125 * c0;
        cxy = -inv12 * (ih[0] + ih[1]);
        cyz = -inv12 * (ih[2] + ih[1]);
        cxz = -inv12 * (ih[0] + ih[2]);
    }
};

// -Laplacian, Mehrstellen 19 points.  FDkernelDel2_4th_Mehr,
// src/pb/FDkernels.cc:448-520 (coefficients :457-463, body :498-513).
// Neighbour groups are summed in T, then scaled in double.
template <typename T>
void del2_mehr(const G& g, const T* v, T* u, long nfunc)
{
    CM c(g.h);
    const long X = g.incx, Y = g.incy;
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                {
                    long i = f * g.sizeg + g.idx(ix, iy, iz);
                    const T* p = v + i;
                    u[i] = (T)(c.c0 * (double)p[0] + c.cz * (double)(T)(p[-1] + p[1])
                               + c.cy * (double)(T)(p[-Y] + p[Y])
                               + c.cx * (double)(T)(p[-X] + p[X])
                               + c.cxz * (double)(T)(p[-X - 1] + p[-X + 1] + p[X - 1] + p[X + 1])
                               + c.cyz * (double)(T)(p[-Y - 1] + p[-Y + 1] + p[Y - 1] + p[Y + 1])
                               + c.cxy
                                     * (double)(T)(p[-X - Y] + p[X - Y] + p[-X + Y] + p[X + Y]));
                }
}

// Mehrstellen right-hand-side operator B.  FDkernelRHS_4th_Mehr1,
// src/pb/FDkernels.cc:523-584: 0.5 v + 1/12 sum(6 neighbours); output grid
// may have another ghost width (rhs_ghosts, 0 = packed).
template <typename T>
void rhs_mehr(const G& g, const T* v, T* rhs, int rg, long nfunc)
{
    const double c0 = 0.5, c1 = inv12;
    const long X = g.incx, Y = g.incy;
    const long incy_r = g.d[2] + 2 * rg;
    const long incx_r = (long)(g.d[1] + 2 * rg) * incy_r;
    const long size_r = (long)(g.d[0] + 2 * rg) * incx_r;
    const long off_r  = rg * incx_r + rg * incy_r + rg;
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                {
                    long i     = f * g.sizeg + g.idx(ix, iy, iz);
                    const T* p = v + i;
                    rhs[f * size_r + off_r + ix * incx_r + iy * incy_r + iz]
                        = (T)(c0 * (double)p[0]
                              + c1 * (double)(T)(p[-X] + p[X] + p[-Y] + p[Y] + p[-1] + p[1]));
                }
}

template <typename T>
int lap_apply(const G& g, int type, const T* in, T* out, long nfunc)
{
    switch (type)
    {
        case 0:
            if (g.ng < 1) return -1;
            del2_mehr(g, in, out, nfunc);
            return 0;
        case 1:
            if (g.ng < 1) return -1;
            del2_2nd(g, in, out, nfunc);
            return 0;
        case 2:
            if (g.ng < 2) return -1;
            del2_4th(g, in, out, nfunc);
            return 0;
        default:
            return -1;
    }
}

// MPaxpy, src/linear_algebra/mputils.cc:233-244: y += (T2)(alpha*(double)x)
template <typename T>
void axpy(long n, double alpha, const T* x, T* y)
{
#pragma omp parallel for
    for (long k = 0; k < n; k++)
        y[k] += static_cast<T>(alpha * static_cast<double>(x[k]));
}

// GridFuncVector::jacobi, src/pb/GridFuncVector.cc:2416-2425:
//   trade(x); w = L x; w -= f; x += -jf * w    (axpys over all sizeg)
template <typename T>
int jacobi(const G& g, const or_grid_t* gg, int type, const T* f, T* x, T* w, double jf,
    long nfunc, const int bc[3])
{
    if (trade(gg, x, nfunc, bc)) return -1;
    if (lap_apply(g, type, x, w, nfunc)) return -1;
    axpy(nfunc * g.sizeg, -1., f, w);
    axpy(nfunc * g.sizeg, -1. * jf, w, x);
    return 0;
}

// Full-weighting restriction.  MGkernelRestrict3D, src/pb/MGkernels.cc:204-283.
template <typename T>
int restrict3d(const G& gf, const T* uf, const G& gc, T* uc, long nfunc)
{
    const long X = gf.incx, Y = gf.incy;
    for (int i = 0; i < 3; i++)
        if (gf.d[i] < 2 * gc.d[i]) return -1;
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < gc.d[0]; ix++)
            for (int iy = 0; iy < gc.d[1]; iy++)
            {
                const T* u0  = uf + f * gf.sizeg + gf.idx(2 * ix, 2 * iy, 0);
                const T* umx = u0 - X;
                const T* upx = u0 + X;
                const T* umy = u0 - Y;
                const T* upy = u0 + Y;
                const T* umxpy = u0 - X + Y;
                const T* upxpy = u0 + X + Y;
                const T* umxmy = u0 - X - Y;
                const T* upxmy = u0 + X - Y;
                T* out         = uc + f * gc.sizeg + gc.idx(ix, iy, 0);
                for (int iz = 0; iz < gc.d[2]; iz++)
                {
                    const int t = 2 * iz;
                    double face = (double)upx[t] + (double)umx[t] + (double)upy[t] + (double)umy[t]
                                  + (double)u0[t - 1] + (double)u0[t + 1];
                    double corner = (double)upxpy[t - 1] + (double)upxpy[t + 1]
                                    + (double)upxmy[t - 1] + (double)upxmy[t + 1]
                                    + (double)umxpy[t - 1] + (double)umxpy[t + 1]
                                    + (double)umxmy[t - 1] + (double)umxmy[t + 1];
                    double edge = (double)upy[t - 1] + (double)upy[t + 1] + (double)umy[t - 1]
                                  + (double)umy[t + 1] + (double)upx[t - 1] + (double)upx[t + 1]
                                  + (double)umx[t - 1] + (double)umx[t + 1] + (double)umxmy[t]
                                  + (double)upxmy[t] + (double)umxpy[t] + (double)upxpy[t];
                    out[iz] = (T)(inv64 * (8. * u0[t] + 4. * face + 2. * edge + corner));
                }
            }
    return 0;
}

// Trilinear prolongation.  MGkernelExtend3D, src/pb/MGkernels.cc:29-201:
// inject coarse (i) -> fine (2i) for i = 0..cdim (first + ghost included,
// :65-81), then fill odd points from injected neighbours in T arithmetic.
template <typename T>
int extend3d(const G& gc, const T* uc, const G& gf, T* uf, long nfunc)
{
    for (int i = 0; i < 3; i++)
    {
        if (gf.d[i] < 2 * gc.d[i]) return -1;
        if (gf.ng < 1 || gc.ng < 1) return -1;
    }
    const long X = gf.incx, Y = gf.incy;
#pragma omp parallel for
    for (long f = 0; f < nfunc; f++)
    {
        T* fd       = uf + f * gf.sizeg;
        const T* cd = uc + f * gc.sizeg;
        for (int ix = 0; ix <= gc.d[0]; ix++)
            for (int iy = 0; iy <= gc.d[1]; iy++)
                for (int iz = 0; iz <= gc.d[2]; iz++)
                    fd[gf.idx(2 * ix, 2 * iy, 2 * iz)] = cd[gc.idx(ix, iy, iz)];
        const int ng = gf.ng;
        for (int ix = ng; ix < gf.d[0] + ng; ix += 2)
        {
            long ifx = X * ix;
            for (int iy = ng; iy < gf.d[1] + ng; iy += 2)
            {
                long ify = ifx + Y * iy;
                for (int iz = ng + 1; iz < gf.d[2] + ng; iz += 2)
                {
                    long i = ify + iz;
                    fd[i]  = (T)(0.5 * (T)(fd[i - 1] + fd[i + 1]));
                }
            }
            for (int iy = ng + 1; iy < gf.d[1] + ng; iy += 2)
            {
                long ify = ifx + Y * iy;
                for (int iz = ng; iz < gf.d[2] + ng; iz += 2)
                {
                    long i = ify + iz;
                    fd[i]  = (T)(0.5 * (T)(fd[i + Y] + fd[i - Y]));
                }
                for (int iz = ng + 1; iz < gf.d[2] + ng; iz += 2)
                {
                    long i = ify + iz;
                    fd[i]  = (T)(0.25 * (T)(fd[i + 1 + Y] + fd[i + 1 - Y] + fd[i - 1 + Y] + fd[i - 1 - Y]));
                }
            }
        }
        for (int ix = ng + 1; ix < gf.d[0] + ng; ix += 2)
        {
            long ifx = X * ix;
            for (int iy = ng; iy < gf.d[1] + ng; iy += 2)
            {
                long ify = ifx + Y * iy;
                for (int iz = ng + 1; iz < gf.d[2] + ng; iz += 2)
                {
                    long i = ify + iz;
                    fd[i]  = (T)(0.25 * (T)(fd[i + X + 1] + fd[i + X - 1] + fd[i - X + 1] + fd[i - X - 1]));
                }
                for (int iz = ng; iz < gf.d[2] + ng; iz += 2)
                {
                    long i = ify + iz;
                    fd[i]  = (T)(0.5 * (T)(fd[i + X] + fd[i - X]));
                }
            }
            for (int iy = ng + 1; iy < gf.d[1] + ng; iy += 2)
            {
                long ify = ifx + Y * iy;
                for (int iz = ng + 1; iz < gf.d[2] + ng; iz += 2)
                {
                    long i = ify + iz;
                    fd[i]  = (T)(0.125
                                * (T)(fd[i + X + Y + 1] + fd[i + X + Y - 1] + fd[i + X - Y + 1]
                                      + fd[i + X - Y - 1] + fd[i - X + Y + 1] + fd[i - X + Y - 1]
                                      + fd[i - X - Y + 1] + fd[i - X - Y - 1]));
                }
                for (int iz = ng; iz < gf.d[2] + ng; iz += 2)
                {
                    long i = ify + iz;
                    fd[i]  = (T)(0.25 * (T)(fd[i + X + Y] + fd[i + X - Y] + fd[i - X + Y] + fd[i - X - Y]));
                }
            }
        }
    }
    return 0;
}

// Cube / general 0-1 mask on interior points.  GridMaskMult applied through
// Map2Masks::apply, src/Map2Masks.cc:26-58 (a mask value of 0 zeroes the
// point, 1 keeps it).
template <typename T>
void mask_apply(const G& g, T* u, const uint8_t* m, long nfunc)
{
    if (!m) return;
#pragma omp parallel for collapse(2)
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                {
                    long p = f * g.size + ((long)ix * g.d[1] + iy) * g.d[2] + iz;
                    if (!m[p]) u[f * g.sizeg + g.idx(ix, iy, iz)] = (T)0;
                }
}

or_grid_t coarse_of(const or_grid_t& g)
{
    // Grid::coarse_grid, src/pb/Grid.cc:214-231: dims halved, same ghosts,
    // same lattice => spacing doubled.
    or_grid_t c = g;
    for (int i = 0; i < 3; i++)
    {
        c.dim[i] = g.dim[i] / 2;
        c.h[i]   = 2. * g.h[i];
    }
    return c;
}

// Preconditioning<T>::mg, src/Preconditioning.cc:155-216.
template <typename T>
struct MGLevels
{
    std::vector<or_grid_t> grids;
    std::vector<double> jf;
    std::vector<std::vector<T>> work, rcoarse, newv;
};

template <typename T>
int mg_rec(MGLevels<T>& L, int max_levels, T* v, const T* f, int lap_type, int level, long nfunc,
    const int bc[3], const uint8_t* const* masks)
{
    const or_grid_t& gg = L.grids[level];
    G g(&gg);
    const int ncycl = (level == max_levels) ? 4 : 2;
    const double jf = L.jf[level];
    T* w            = L.work[level].data();
    const uint8_t* mk = masks ? masks[level] : nullptr;
    for (int it = 0; it < ncycl; it++)
    {
        if (jacobi(g, &gg, lap_type, f, v, w, jf, nfunc, bc)) return -1;
        mask_apply(g, v, mk, nfunc);
    }
    if (level == max_levels) return 0;

    mask_apply(g, w, mk, nfunc);
    const or_grid_t& gcg = L.grids[level + 1];
    G gc(&gcg);
    T* rc = L.rcoarse[level].data();
    trade(&gg, w, nfunc, bc);
    restrict3d(g, w, gc, rc, nfunc);
    mask_apply(gc, rc, masks ? masks[level + 1] : nullptr, nfunc);

    T* nv = L.newv[level].data();
    std::fill(nv, nv + nfunc * gc.sizeg, (T)0);
    if (mg_rec(L, max_levels, nv, rc, 1, level + 1, nfunc, bc, masks)) return -1;

    trade(&gcg, nv, nfunc, bc);
    extend3d(gc, nv, g, w, nfunc);
    mask_apply(g, w, mk, nfunc);

    axpy(nfunc * g.sizeg, -1., w, v);

    for (int it = 0; it < 2; it++)
    {
        if (jacobi(g, &gg, lap_type, f, v, w, jf, nfunc, bc)) return -1;
        mask_apply(g, v, mk, nfunc);
    }
    if (bc[0] != 1 || bc[1] != 1 || bc[2] != 1) trade(&gg, v, nfunc, bc);
    return 0;
}

double diag_el(int type, const double h[3])
{
    const double s = 1. / (h[0] * h[0]) + 1. / (h[1] * h[1]) + 1. / (h[2] * h[2]);
    switch (type)
    {
        case 0: return (4. / 3.) * s;  // Laph4M, src/pb/Laph4M.h:29-31
        case 1: return 2. * s;         // Laph2,  src/pb/Laph2.h:32-33
        case 2: return 2.5 * s;        // Laph4,  src/pb/Laph4.h:35-38
        default: return 0.;
    }
}

double jacobi_factor(int type, const double h[3])
{
    const double inv = 1. / diag_el(type, h);
    switch (type)
    {
        case 0: return inv;            // src/pb/Laph4M.h:80
        case 1: return 6. * inv / 7.;  // src/pb/Laph2.h:95
        case 2: return inv / 1.5;      // src/pb/Laph4.h:106
        default: return 0.;
    }
}

template <typename T>
int mg_levels_setup(MGLevels<T>& L, const or_grid_t* g0, int lap_type, int max_levels, long nfunc)
{
    // Preconditioning ctor + setup, src/Preconditioning.cc:15-28, 98-151:
    // level 0 uses lap_type's factor, coarse levels use Laph2 (type 1).
    if (lap_type < 0 || lap_type > 2) return -1;
    L.grids.push_back(*g0);
    L.jf.push_back(jacobi_factor(lap_type, g0->h));
    for (int l = 1; l <= max_levels; l++)
    {
        const or_grid_t& p = L.grids.back();
        for (int i = 0; i < 3; i++)
            if (p.dim[i] % 2) return -1;
        or_grid_t c = coarse_of(p);
        L.grids.push_back(c);
        L.jf.push_back(jacobi_factor(1, c.h));
    }
    for (int l = 0; l <= max_levels; l++)
    {
        G g(&L.grids[l]);
        L.work.emplace_back(nfunc * g.sizeg, (T)0);
        if (l < max_levels)
        {
            G gc(&L.grids[l + 1]);
            L.rcoarse.emplace_back(nfunc * gc.sizeg, (T)0);
            L.newv.emplace_back(nfunc * gc.sizeg, (T)0);
        }
    }
    return 0;
}

template <typename T>
int mg(const or_grid_t* g0, int lap_type, int max_levels, T* v, const T* f, long nfunc,
    const int bc[3], const uint8_t* const* masks)
{
    MGLevels<T> L;
    if (mg_levels_setup(L, g0, lap_type, max_levels, nfunc)) return -1;
    return mg_rec(L, max_levels, v, f, lap_type, 0, nfunc, bc, masks);
}

// OrbitalsPreconditioning<T>::precond_mg, src/OrbitalsPreconditioning.cc:87-117:
// x = 0; f = (T)res with zero ghosts; x += gamma*f; mg(x, f); out = x.
template <typename T>
int precond(const or_grid_t* g0, int lap_type, int max_levels, double gamma, const double* res,
    double* out, long ld, long nfunc, const int bc[3], const uint8_t* const* masks)
{
    G g(g0);
    std::vector<T> x(nfunc * g.sizeg, (T)0), f(nfunc * g.sizeg, (T)0);
    for (long fn = 0; fn < nfunc; fn++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                {
                    long p = ((long)ix * g.d[1] + iy) * g.d[2] + iz;
                    f[fn * g.sizeg + g.idx(ix, iy, iz)] = (T)res[fn * ld + p];
                }
    axpy(nfunc * g.sizeg, gamma, f.data(), x.data());
    if (mg(g0, lap_type, max_levels, x.data(), f.data(), nfunc, bc, masks)) return -1;
    for (long fn = 0; fn < nfunc; fn++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                {
                    long p          = ((long)ix * g.d[1] + iy) * g.d[2] + iz;
                    out[fn * ld + p] = (double)x[fn * g.sizeg + g.idx(ix, iy, iz)];
                }
    return 0;
}

// ===========================================================================
// Box family.  Box j, level l: edge e = edge_j >> l, origin o = origin_j >> l
// on the periodic grid gdim >> l; data packed x-slowest, zero outside the box.
// Every op equals the grid op on the box embedded in the periodic grid with a
// 0/1 cube mask applied as in Preconditioning::mg.
struct Boxes
{
    const or_boxes_t* b;
    int L;  // number of levels with precomputed offsets
    std::vector<std::vector<long>> off;  // off[l][j], off[l][nbox] = total
    Boxes(const or_boxes_t* bb, int levels) : b(bb), L(levels)
    {
        off.resize(levels);
        for (int l = 0; l < levels; l++)
        {
            off[l].resize(b->nbox + 1);
            long s = 0;
            for (int j = 0; j < b->nbox; j++)
            {
                off[l][j] = s;
                long e    = b->edge[j] >> l;
                s += e * e * e;
            }
            off[l][b->nbox] = s;
        }
    }
    int e(int l, int j) const { return b->edge[j] >> l; }
    int o(int l, int j, int d) const { return b->origin[3 * j + d] >> l; }
    int gd(int l, int d) const { return b->gdim[d] >> l; }
    double h(int l, int d) const { return b->h[d] * (double)(1 << l); }
};

template <typename T>
inline T bget(const T* p, int e, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= e || y >= e || z >= e) return (T)0;
    return p[((long)x * e + y) * e + z];
}

template <typename T>
void b_lap_box(int type, const double h[3], const T* in, T* out, int e, const double* vglob,
    const Boxes* B, int j)
{
    if (type == 2)
    {
        C4 c(h);
        for (int x = 0; x < e; x++)
            for (int y = 0; y < e; y++)
                for (int z = 0; z < e; z++)
                {
                    double diag = c.c0;
                    if (vglob)
                    {
                        int gx   = wrap(B->o(0, j, 0) + x, B->gd(0, 0));
                        int gy   = wrap(B->o(0, j, 1) + y, B->gd(0, 1));
                        int gz   = wrap(B->o(0, j, 2) + z, B->gd(0, 2));
                        diag = c.c0 + vglob[((long)gx * B->gd(0, 1) + gy) * B->gd(0, 2) + gz];
                    }
                    double r = diag * (double)bget(in, e, x, y, z)
                               + c.c1x * ((double)bget(in, e, x - 1, y, z) + (double)bget(in, e, x + 1, y, z))
                               + c.c1y * ((double)bget(in, e, x, y - 1, z) + (double)bget(in, e, x, y + 1, z))
                               + c.c1z * ((double)bget(in, e, x, y, z - 1) + (double)bget(in, e, x, y, z + 1))
                               + c.c2x * ((double)bget(in, e, x - 2, y, z) + (double)bget(in, e, x + 2, y, z))
                               + c.c2y * ((double)bget(in, e, x, y - 2, z) + (double)bget(in, e, x, y + 2, z))
                               + c.c2z * ((double)bget(in, e, x, y, z - 2) + (double)bget(in, e, x, y, z + 2));
                    out[((long)x * e + y) * e + z] = (T)r;
                }
    }
    else if (type == 1)
    {
        const double c1x = -1. / (h[0] * h[0]), c1y = -1. / (h[1] * h[1]), c1z = -1. / (h[2] * h[2]);
        const double c0 = -2. * (c1x + c1y + c1z);
        for (int x = 0; x < e; x++)
            for (int y = 0; y < e; y++)
                for (int z = 0; z < e; z++)
                {
                    double r = c0 * bget(in, e, x, y, z)
                               + c1x * (T)(bget(in, e, x - 1, y, z) + bget(in, e, x + 1, y, z))
                               + c1y * (T)(bget(in, e, x, y - 1, z) + bget(in, e, x, y + 1, z))
                               + c1z * (T)(bget(in, e, x, y, z - 1) + bget(in, e, x, y, z + 1));
                    out[((long)x * e + y) * e + z] = (T)r;
                }
    }
    else  // Mehrstellen
    {
        CM c(h);
        for (int x = 0; x < e; x++)
            for (int y = 0; y < e; y++)
                for (int z = 0; z < e; z++)
                {
#define V(dx, dy, dz) bget(in, e, x + (dx), y + (dy), z + (dz))
                    double r = c.c0 * (double)V(0, 0, 0) + c.cz * (double)(T)(V(0, 0, -1) + V(0, 0, 1))
                               + c.cy * (double)(T)(V(0, -1, 0) + V(0, 1, 0))
                               + c.cx * (double)(T)(V(-1, 0, 0) + V(1, 0, 0))
                               + c.cxz * (double)(T)(V(-1, 0, -1) + V(-1, 0, 1) + V(1, 0, -1) + V(1, 0, 1))
                               + c.cyz * (double)(T)(V(0, -1, -1) + V(0, -1, 1) + V(0, 1, -1) + V(0, 1, 1))
                               + c.cxy * (double)(T)(V(-1, -1, 0) + V(1, -1, 0) + V(-1, 1, 0) + V(1, 1, 0));
#undef V
                    out[((long)x * e + y) * e + z] = (T)r;
                }
    }
}

template <typename T>
int b_lap(const or_boxes_t* bb, int level, int type, const T* in, T* out)
{
    if (type < 0 || type > 2) return -1;
    Boxes B(bb, level + 1);
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < bb->nbox; j++)
    {
        double h[3] = { B.h(level, 0), B.h(level, 1), B.h(level, 2) };
        long o      = B.off[level][j];
        b_lap_box(type, h, in + o, out + o, B.e(level, j), (const double*)nullptr, &B, j);
    }
    return 0;
}

template <typename T>
int b_hloc(const or_boxes_t* bb, const double* vglob, const T* in, T* out)
{
    Boxes B(bb, 1);
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < bb->nbox; j++)
    {
        long o = B.off[0][j];
        b_lap_box(2, bb->h, in + o, out + o, B.e(0, j), vglob, &B, j);
    }
    return 0;
}

template <typename T>
int b_jacobi(const or_boxes_t* bb, int level, int type, const T* f, T* x, T* w, double jf)
{
    if (b_lap(bb, level, type, x, w)) return -1;
    Boxes B(bb, level + 1);
    long n = B.off[level][bb->nbox];
    axpy(n, -1., f, w);
    axpy(n, -1. * jf, w, x);
    return 0;
}

template <typename T>
int b_restrict(const or_boxes_t* bb, int level, const T* fine, T* coarse)
{
    Boxes B(bb, level + 2);
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < bb->nbox; j++)
    {
        const T* uf = fine + B.off[level][j];
        T* uc       = coarse + B.off[level + 1][j];
        const int ef = B.e(level, j), ec = B.e(level + 1, j);
        for (int x = 0; x < ec; x++)
            for (int y = 0; y < ec; y++)
                for (int z = 0; z < ec; z++)
                {
                    const int X = 2 * x, Y = 2 * y, Z = 2 * z;
#define F(dx, dy, dz) (double)bget(uf, ef, X + (dx), Y + (dy), Z + (dz))
                    double face = F(1, 0, 0) + F(-1, 0, 0) + F(0, 1, 0) + F(0, -1, 0) + F(0, 0, -1) + F(0, 0, 1);
                    double corner = F(1, 1, -1) + F(1, 1, 1) + F(1, -1, -1) + F(1, -1, 1) + F(-1, 1, -1)
                                    + F(-1, 1, 1) + F(-1, -1, -1) + F(-1, -1, 1);
                    double edge = F(0, 1, -1) + F(0, 1, 1) + F(0, -1, -1) + F(0, -1, 1) + F(1, 0, -1)
                                  + F(1, 0, 1) + F(-1, 0, -1) + F(-1, 0, 1) + F(-1, -1, 0) + F(1, -1, 0)
                                  + F(-1, 1, 0) + F(1, 1, 0);
                    uc[((long)x * ec + y) * ec + z]
                        = (T)(inv64 * (8. * F(0, 0, 0) + 4. * face + 2. * edge + corner));
#undef F
                }
    }
    return 0;
}

template <typename T>
int b_prolong(const or_boxes_t* bb, int level, const T* coarse, T* fine)
{
    // fine box at `level`, coarse box at level+1.  Injected value at even
    // fine coordinates is the coarse value (0 outside the coarse box).
    Boxes B(bb, level + 2);
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < bb->nbox; j++)
    {
        const T* uc = coarse + B.off[level + 1][j];
        T* uf       = fine + B.off[level][j];
        const int ef = B.e(level, j), ec = B.e(level + 1, j);
        auto I = [&](int x, int y, int z) -> T { return bget(uc, ec, x / 2, y / 2, z / 2); };
        for (int x = 0; x < ef; x++)
            for (int y = 0; y < ef; y++)
                for (int z = 0; z < ef; z++)
                {
                    const int ox = x & 1, oy = y & 1, oz = z & 1;
                    T r;
                    if (!ox && !oy && !oz)
                        r = I(x, y, z);
                    else if (!ox && !oy && oz)
                        r = (T)(0.5 * (T)(I(x, y, z - 1) + I(x, y, z + 1)));
                    else if (!ox && oy && !oz)
                        r = (T)(0.5 * (T)(I(x, y + 1, z) + I(x, y - 1, z)));
                    else if (!ox && oy && oz)
                        r = (T)(0.25 * (T)(I(x, y + 1, z + 1) + I(x, y - 1, z + 1) + I(x, y + 1, z - 1)
                                           + I(x, y - 1, z - 1)));
                    else if (ox && !oy && oz)
                        r = (T)(0.25 * (T)(I(x + 1, y, z + 1) + I(x + 1, y, z - 1) + I(x - 1, y, z + 1)
                                           + I(x - 1, y, z - 1)));
                    else if (ox && !oy && !oz)
                        r = (T)(0.5 * (T)(I(x + 1, y, z) + I(x - 1, y, z)));
                    else if (ox && oy && oz)
                        r = (T)(0.125 * (T)(I(x + 1, y + 1, z + 1) + I(x + 1, y + 1, z - 1)
                                            + I(x + 1, y - 1, z + 1) + I(x + 1, y - 1, z - 1)
                                            + I(x - 1, y + 1, z + 1) + I(x - 1, y + 1, z - 1)
                                            + I(x - 1, y - 1, z + 1) + I(x - 1, y - 1, z - 1)));
                    else  // ox && oy && !oz
                        r = (T)(0.25 * (T)(I(x + 1, y + 1, z) + I(x + 1, y - 1, z) + I(x - 1, y + 1, z)
                                           + I(x - 1, y - 1, z)));
                    uf[((long)x * ef + y) * ef + z] = r;
                }
    }
    return 0;
}

template <typename T>
int b_mg_rec(const or_boxes_t* bb, const Boxes& B, int max_levels, std::vector<std::vector<T>>& work,
    std::vector<std::vector<T>>& rc, std::vector<std::vector<T>>& nv, T* v, const T* f, int type,
    int level)
{
    // Preconditioning<T>::mg, src/Preconditioning.cc:155-216, masks implicit.
    const int ncycl = (level == max_levels) ? 4 : 2;
    double h[3]     = { B.h(level, 0), B.h(level, 1), B.h(level, 2) };
    const double jf = jacobi_factor(level == 0 ? type : 1, h);
    T* w            = work[level].data();
    for (int it = 0; it < ncycl; it++)
        if (b_jacobi(bb, level, type, f, v, w, jf)) return -1;
    if (level == max_levels) return 0;
    b_restrict(bb, level, w, rc[level].data());
    std::fill(nv[level].begin(), nv[level].end(), (T)0);
    if (b_mg_rec(bb, B, max_levels, work, rc, nv, nv[level].data(), rc[level].data(), 1, level + 1))
        return -1;
    b_prolong(bb, level, nv[level].data(), w);
    axpy(B.off[level][bb->nbox], -1., w, v);
    for (int it = 0; it < 2; it++)
        if (b_jacobi(bb, level, type, f, v, w, jf)) return -1;
    return 0;
}

int b_check_levels(const or_boxes_t* bb, int max_levels)
{
    const int m = 1 << max_levels;
    for (int j = 0; j < bb->nbox; j++)
    {
        if (bb->edge[j] % m || bb->edge[j] < m) return -1;
        for (int d = 0; d < 3; d++)
        {
            if (bb->origin[3 * j + d] % m) return -1;
            if (2 * bb->edge[j] + 4 > bb->gdim[d]) return -1;
        }
    }
    for (int d = 0; d < 3; d++)
        if (bb->gdim[d] % m) return -1;
    return 0;
}

template <typename T>
int b_mg(const or_boxes_t* bb, int type, int max_levels, T* v, const T* f)
{
    if (type < 0 || type > 2 || b_check_levels(bb, max_levels)) return -1;
    Boxes B(bb, max_levels + 1);
    std::vector<std::vector<T>> work, rc, nv;
    for (int l = 0; l <= max_levels; l++)
    {
        work.emplace_back(B.off[l][bb->nbox], (T)0);
        if (l < max_levels)
        {
            rc.emplace_back(B.off[l + 1][bb->nbox], (T)0);
            nv.emplace_back(B.off[l + 1][bb->nbox], (T)0);
        }
    }
    return b_mg_rec(bb, B, max_levels, work, rc, nv, v, f, type, 0);
}

template <typename T>
int b_precond(const or_boxes_t* bb, int type, int max_levels, double gamma, const double* res, double* out)
{
    Boxes B(bb, 1);
    const long n = B.off[0][bb->nbox];
    std::vector<T> x(n, (T)0), f(n);
    for (long i = 0; i < n; i++)
        f[i] = (T)res[i];
    axpy(n, gamma, f.data(), x.data());
    if (b_mg(bb, type, max_levels, x.data(), f.data())) return -1;
    for (long i = 0; i < n; i++)
        out[i] = (double)x[i];
    return 0;
}

// Periodic overlap of boxes i and j along direction d, in i-local coords
// [a, b); sh = offset to add to an i-local coordinate to get a j-local one.
bool overlap1d(int oi, int ei, int oj, int ej, int G, int* a, int* b, int* sh)
{
    int s = wrap(oj - oi, G);
    // candidate j ranges relative to oi: [s, s+ej) and [s-G, s-G+ej)
    int a1 = std::max(0, s), b1 = std::min(ei, s + ej);
    if (a1 < b1)
    {
        *a  = a1;
        *b  = b1;
        *sh = -s;
        return true;
    }
    int a2 = std::max(0, s - G), b2 = std::min(ei, s - G + ej);
    if (a2 < b2)
    {
        *a  = a2;
        *b  = b2;
        *sh = G - s;
        return true;
    }
    return false;
}

struct Ovl
{
    int a[3], b[3], sh[3];
};

bool box_overlap(const or_boxes_t* bb, int i, int j, Ovl* o)
{
    for (int d = 0; d < 3; d++)
        if (!overlap1d(bb->origin[3 * i + d], bb->edge[i], bb->origin[3 * j + d], bb->edge[j],
                bb->gdim[d], &o->a[d], &o->b[d], &o->sh[d]))
            return false;
    return true;
}

// Stencil part of (H psi_j)(r) at a point r outside box j (j-local coords):
// only the axis-aligned 4th-order legs that land inside the box contribute
// (the potential term vanishes there because psi_j(r) = 0).
double skin_value(const double* pj, int ej, const C4& c, int x, int y, int z)
{
    int cnt_out = (x < 0 || x >= ej) + (y < 0 || y >= ej) + (z < 0 || z >= ej);
    if (cnt_out != 1) return 0.;
    double v = 0.;
    auto leg = [&](int ax, int ay, int az, double cf) {
        if (ax < 0 || ay < 0 || az < 0 || ax >= ej || ay >= ej || az >= ej) return;
        v += cf * pj[((long)ax * ej + ay) * ej + az];
    };
    leg(x - 1, y, z, c.c1x);
    leg(x + 1, y, z, c.c1x);
    leg(x - 2, y, z, c.c2x);
    leg(x + 2, y, z, c.c2x);
    leg(x, y - 1, z, c.c1y);
    leg(x, y + 1, z, c.c1y);
    leg(x, y - 2, z, c.c2y);
    leg(x, y + 2, z, c.c2y);
    leg(x, y, z - 1, c.c1z);
    leg(x, y, z + 1, c.c1z);
    leg(x, y, z - 2, c.c2z);
    leg(x, y, z + 2, c.c2z);
    return v;
}

// sum over r in box i, r outside box j, of psi_i(r) (H psi_j)(r);
// sh maps i-local -> j-local coordinates.  Only points one axis away from box
// j (the 12 slabs of thickness 1 at distance 1 and 2 from its faces) reach it.
double skin_sum(const double* pi, int ei, const double* pj, int ej, const int sh[3], const C4& c)
{
    double acc = 0.;
    int lo[3], hi[3];
    for (int k = 0; k < 3; k++)
    {
        lo[k] = std::max(0, sh[k]);
        hi[k] = std::min(ej, sh[k] + ei);
    }
    for (int slab = 0; slab < 12; slab++)
    {
        const int d = slab >> 2, side = (slab >> 1) & 1, layer = (slab & 1) + 1;
        const int Xd = side == 0 ? -layer : ej - 1 + layer;
        const int xd = Xd - sh[d];
        if (xd < 0 || xd >= ei) continue;
        const int u = (d + 1) % 3, w = (d + 2) % 3;
        for (int a = lo[u]; a < hi[u]; a++)
            for (int b = lo[w]; b < hi[w]; b++)
            {
                int X[3];
                X[d] = Xd;
                X[u] = a;
                X[w] = b;
                const double hv = skin_value(pj, ej, c, X[0], X[1], X[2]);
                acc += pi[((long)(X[0] - sh[0]) * ei + (X[1] - sh[1])) * ei + (X[2] - sh[2])] * hv;
            }
    }
    return acc;
}

// Periodic shift (i-local -> j-local) for boxes whose extended regions (2
// points around box j) meet box i; false if they do not interact.
bool ext_shift(const or_boxes_t* b, int i, int j, int sh[3])
{
    for (int d = 0; d < 3; d++)
    {
        const int G = b->gdim[d];
        const int s = wrap(b->origin[3 * j + d] - b->origin[3 * i + d], G);
        // j-local X = x - s (mod G); pick the image closest to the box
        int best = 1 << 30, bsh = 0;
        for (int img = -1; img <= 1; img++)
        {
            const int shift = -s + img * G;  // X = x + shift
            const int lo = shift, hi = b->edge[i] - 1 + shift;  // range of X
            // distance of [lo,hi] from [-2, ej+1]
            int dist = 0;
            if (hi < -2) dist = -2 - hi;
            else if (lo > b->edge[j] + 1) dist = lo - (b->edge[j] + 1);
            if (dist < best)
            {
                best = dist;
                bsh  = shift;
            }
        }
        if (best > 0) return false;
        sh[d] = bsh;
    }
    return true;
}

}  // namespace

// ===========================================================================
extern "C" {

int or_g_trade_f32(const or_grid_t* g, float* u, long nfunc, const int bc[3]) { return trade(g, u, nfunc, bc); }
int or_g_trade_f64(const or_grid_t* g, double* u, long nfunc, const int bc[3]) { return trade(g, u, nfunc, bc); }

int or_g_lap_f32(const or_grid_t* g, int t, const float* in, float* out, long n) { return lap_apply(G(g), t, in, out, n); }
int or_g_lap_f64(const or_grid_t* g, int t, const double* in, double* out, long n) { return lap_apply(G(g), t, in, out, n); }

int or_g_rhs_mehr_f32(const or_grid_t* g, const float* in, float* out, int rg, long n)
{
    if (g->ghosts < 1) return -1;
    rhs_mehr(G(g), in, out, rg, n);
    return 0;
}
int or_g_rhs_mehr_f64(const or_grid_t* g, const double* in, double* out, int rg, long n)
{
    if (g->ghosts < 1) return -1;
    rhs_mehr(G(g), in, out, rg, n);
    return 0;
}

int or_g_lap4_withpot_f32(const or_grid_t* g, const float* in, const double* pot, float* out, long ld, long n)
{
    if (g->ghosts < 2) return -1;
    del2_4th_withpot(G(g), in, pot, out, ld, n);
    return 0;
}
int or_g_lap4_withpot_f64(const or_grid_t* g, const double* in, const double* pot, double* out, long ld, long n)
{
    if (g->ghosts < 2) return -1;
    del2_4th_withpot(G(g), in, pot, out, ld, n);
    return 0;
}

// GridFuncVector::pointwiseProduct, src/pb/GridFuncVector.cc:90-136: over all
// sizeg points (ghosts included): out = (T)(a*b).
int or_g_pointwise_f32(const or_grid_t* gg, const float* a, const double* b, float* out, long n)
{
    G g(gg);
    for (long f = 0; f < n; f++)
        for (long i = 0; i < g.sizeg; i++)
            out[f * g.sizeg + i] = (float)(a[f * g.sizeg + i] * b[i]);
    return 0;
}
int or_g_pointwise_f64(const or_grid_t* gg, const double* a, const double* b, double* out, long n)
{
    G g(gg);
    for (long f = 0; f < n; f++)
        for (long i = 0; i < g.sizeg; i++)
            out[f * g.sizeg + i] = (double)(a[f * g.sizeg + i] * b[i]);
    return 0;
}

// Hamiltonian<T>::applyLocal(ncolors, phi, hphi), src/Hamiltonian.cc:86-159.
// phi: ghosted block (traded here), vtot packed (numpt), hphi packed (ld).
int or_g_hloc_f64(const or_grid_t* gg, int lap_type, const double* phi_in, const double* vtot,
    double* hphi, long ld, long nfunc, const int bc[3])
{
    G g(gg);
    std::vector<double> phi(phi_in, phi_in + nfunc * g.sizeg);
    trade(gg, phi.data(), nfunc, bc);
    if (lap_type == 2)
    {
        if (g.ng < 2) return -1;
        del2_4th_withpot(g, phi.data(), vtot, hphi, ld, nfunc);
        return 0;
    }
    if (lap_type != 0) return -1;
    // Mehrstellen: w1 = phi*V (ghosts incl.), work1 = B w1, w1 = L_M phi,
    // work1 += w1, hphi = interior(work1)   (src/Hamiltonian.cc:106-131)
    std::vector<double> vg(g.sizeg, 0.);
    for (int ix = 0; ix < g.d[0]; ix++)
        for (int iy = 0; iy < g.d[1]; iy++)
            for (int iz = 0; iz < g.d[2]; iz++)
                vg[g.idx(ix, iy, iz)] = vtot[((long)ix * g.d[1] + iy) * g.d[2] + iz];
    trade(gg, vg.data(), 1, bc);
    std::vector<double> w1(nfunc * g.sizeg), work1(nfunc * g.sizeg, 0.);
    or_g_pointwise_f64(gg, phi.data(), vg.data(), w1.data(), nfunc);
    rhs_mehr(g, w1.data(), work1.data(), g.ng, nfunc);
    del2_mehr(g, phi.data(), w1.data(), nfunc);
    axpy(nfunc * g.sizeg, 1., w1.data(), work1.data());
    for (long f = 0; f < nfunc; f++)
        for (int ix = 0; ix < g.d[0]; ix++)
            for (int iy = 0; iy < g.d[1]; iy++)
                for (int iz = 0; iz < g.d[2]; iz++)
                    hphi[f * ld + ((long)ix * g.d[1] + iy) * g.d[2] + iz]
                        = work1[f * g.sizeg + g.idx(ix, iy, iz)];
    return 0;
}

int or_axpy_f32(long n, double a, const float* x, float* y)
{
    axpy(n, a, x, y);
    return 0;
}
int or_axpy_f64(long n, double a, const double* x, double* y)
{
    axpy(n, a, x, y);
    return 0;
}

int or_g_jacobi_f32(const or_grid_t* g, int t, const float* f, float* x, float* w, double jf, long n, const int bc[3])
{
    return jacobi(G(g), g, t, f, x, w, jf, n, bc);
}
int or_g_jacobi_f64(const or_grid_t* g, int t, const double* f, double* x, double* w, double jf, long n, const int bc[3])
{
    return jacobi(G(g), g, t, f, x, w, jf, n, bc);
}

int or_g_restrict_f32(const or_grid_t* gf, const float* uf, const or_grid_t* gc, float* uc, long n)
{
    return restrict3d(G(gf), uf, G(gc), uc, n);
}
int or_g_restrict_f64(const or_grid_t* gf, const double* uf, const or_grid_t* gc, double* uc, long n)
{
    return restrict3d(G(gf), uf, G(gc), uc, n);
}
int or_g_extend_f32(const or_grid_t* gc, const float* uc, const or_grid_t* gf, float* uf, long n)
{
    return extend3d(G(gc), uc, G(gf), uf, n);
}
int or_g_extend_f64(const or_grid_t* gc, const double* uc, const or_grid_t* gf, double* uf, long n)
{
    return extend3d(G(gc), uc, G(gf), uf, n);
}

int or_g_mask_f32(const or_grid_t* g, float* u, const uint8_t* m, long n)
{
    mask_apply(G(g), u, m, n);
    return 0;
}
int or_g_mask_f64(const or_grid_t* g, double* u, const uint8_t* m, long n)
{
    mask_apply(G(g), u, m, n);
    return 0;
}

double or_diag_el(int t, const double h[3]) { return diag_el(t, h); }
double or_jacobi_factor(int t, const double h[3]) { return jacobi_factor(t, h); }

int or_g_mg_f32(const or_grid_t* g0, int t, int ml, float* v, const float* f, long n, const int bc[3],
    const uint8_t* const* masks)
{
    return mg(g0, t, ml, v, f, n, bc, masks);
}
int or_g_mg_f64(const or_grid_t* g0, int t, int ml, double* v, const double* f, long n, const int bc[3],
    const uint8_t* const* masks)
{
    return mg(g0, t, ml, v, f, n, bc, masks);
}

int or_g_precond_f32(const or_grid_t* g0, int t, int ml, double gamma, const double* res, double* out,
    long ld, long n, const int bc[3], const uint8_t* const* masks)
{
    return precond<float>(g0, t, ml, gamma, res, out, ld, n, bc, masks);
}
int or_g_precond_f64(const or_grid_t* g0, int t, int ml, double gamma, const double* res, double* out,
    long ld, long n, const int bc[3], const uint8_t* const* masks)
{
    return precond<double>(g0, t, ml, gamma, res, out, ld, n, bc, masks);
}

// OrbitalsPreconditioning<T>::setGamma, src/OrbitalsPreconditioning.cc:120-145,
// with the fine operator's invDiagEl (Laph4: 1/(2.5 sum 1/h^2)).
double or_gamma(const double h[3], int mg_levels, double vmax, double small_eig)
{
    double g = 1. / diag_el(2, h);
    for (int ln = 0; ln <= mg_levels; ln++)
        g *= 4.;
    return 1.0 / (2.0 / g + std::fabs(vmax - small_eig));
}

// LocalMatrices::syrk per slab, then fillUpperWithLower and scal(vel):
// LocGridOrbitals::getLocalOverlap, src/LocGridOrbitals.cc:1504-1530;
// LocalMatrices.cc:210-247 (DSYRK 'l','t').  psi column-major (ld rows).
int or_local_gram(const double* psi, long ld, long lnp, int n, int subdivx, double vel, double* ss)
{
#pragma omp parallel for collapse(2)
    for (int s = 0; s < subdivx; s++)
        for (int i = 0; i < n; i++)
            for (int j = 0; j <= i; j++)
            {
                const double* a = psi + (long)i * ld + s * lnp;
                const double* b = psi + (long)j * ld + s * lnp;
                double acc      = 0.;
                for (long k = 0; k < lnp; k++)
                    acc += a[k] * b[k];
                ss[(long)s * n * n + (long)j * n + i] = acc * vel;
                ss[(long)s * n * n + (long)i * n + j] = acc * vel;
            }
    return 0;
}

// LocGridOrbitals::computeLocalProduct, src/LocGridOrbitals.cc:1554-1604
// (gemm 't','n' per slab, LocalMatrices.cc:300-338), then scal(vel).
int or_local_product(const double* A, const double* Bm, long ld, long lnp, int n, int subdivx, double vel,
    double* ss)
{
#pragma omp parallel for collapse(2)
    for (int s = 0; s < subdivx; s++)
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
            {
                const double* a = A + (long)i * ld + s * lnp;
                const double* b = Bm + (long)j * ld + s * lnp;
                double acc      = 0.;
                for (long k = 0; k < lnp; k++)
                    acc += a[k] * b[k];
                ss[(long)s * n * n + (long)j * n + i] = acc * vel;
            }
    return 0;
}

// LocalMatrices2DistMatrix::convert, src/LocalMatrices2DistMatrix.cc:21-78:
// entries with |x| <= tol are dropped, same-pair slab entries summed, then
// pushed into the global matrix (here a dense column-major numst x numst sum).
int or_local_to_global(const double* ss, const int* gids, int subdivx, int m, int numst, double tol,
    double* global)
{
    for (int ic = 0; ic < m; ic++)
        for (int jc = 0; jc < m; jc++)
        {
            int pst_old = -1;
            std::vector<double> val;
            std::vector<int> ist;
            for (int s = 0; s < subdivx; s++)
            {
                const int st1 = gids[s * m + ic];
                if (st1 == -1) continue;
                const int st2 = gids[s * m + jc];
                if (st2 == -1) continue;
                const int pst  = st2 * numst + st1;
                const double t = ss[(long)s * m * m + (long)jc * m + ic];
                if (std::fabs(t) > tol)
                {
                    if (pst == pst_old)
                        val.back() += t;
                    else
                    {
                        val.push_back(t);
                        ist.push_back(st1);
                        ist.push_back(st2);
                        pst_old = pst;
                    }
                }
            }
            for (size_t k = 0; k < val.size(); k++)
                global[(long)ist[2 * k + 1] * numst + ist[2 * k]] += val[k];
        }
    return 0;
}

// MGmol::computeResidualUsingHPhi, src/MGmol.cc:1227-1291 (non-Mehrstellen):
// res = psi*theta (multiplyByMatrix / MPgemmNN) ; res -= hphi.
int or_residual(const double* psi, const double* hphi, const double* theta, long ld, long numpt, int n,
    double* res)
{
#pragma omp parallel for
    for (int j = 0; j < n; j++)
        for (long k = 0; k < numpt; k++)
        {
            double acc = 0.;
            for (int i = 0; i < n; i++)
                acc += psi[(long)i * ld + k] * theta[(long)j * n + i];
            res[(long)j * ld + k] = acc;
            res[(long)j * ld + k] += -1. * hphi[(long)j * ld + k];
        }
    return 0;
}

// ------------------------------- box family -------------------------------
long or_b_size(const or_boxes_t* b, int level)
{
    Boxes B(b, level + 1);
    return B.off[level][b->nbox];
}
long or_b_offset(const or_boxes_t* b, int level, int j)
{
    Boxes B(b, level + 1);
    return B.off[level][j];
}

int or_b_lap_f32(const or_boxes_t* b, int l, int t, const float* in, float* out) { return b_lap(b, l, t, in, out); }
int or_b_lap_f64(const or_boxes_t* b, int l, int t, const double* in, double* out) { return b_lap(b, l, t, in, out); }
int or_b_hloc_f32(const or_boxes_t* b, const double* v, const float* in, float* out) { return b_hloc(b, v, in, out); }
int or_b_hloc_f64(const or_boxes_t* b, const double* v, const double* in, double* out) { return b_hloc(b, v, in, out); }
int or_b_jacobi_f32(const or_boxes_t* b, int l, int t, const float* f, float* x, float* w, double jf)
{
    return b_jacobi(b, l, t, f, x, w, jf);
}
int or_b_jacobi_f64(const or_boxes_t* b, int l, int t, const double* f, double* x, double* w, double jf)
{
    return b_jacobi(b, l, t, f, x, w, jf);
}
int or_b_restrict_f32(const or_boxes_t* b, int l, const float* fi, float* co) { return b_restrict(b, l, fi, co); }
int or_b_restrict_f64(const or_boxes_t* b, int l, const double* fi, double* co) { return b_restrict(b, l, fi, co); }
int or_b_prolong_f32(const or_boxes_t* b, int l, const float* co, float* fi) { return b_prolong(b, l, co, fi); }
int or_b_prolong_f64(const or_boxes_t* b, int l, const double* co, double* fi) { return b_prolong(b, l, co, fi); }
int or_b_mg_f32(const or_boxes_t* b, int t, int ml, float* v, const float* f) { return b_mg(b, t, ml, v, f); }
int or_b_mg_f64(const or_boxes_t* b, int t, int ml, double* v, const double* f) { return b_mg(b, t, ml, v, f); }
int or_b_precond_f32(const or_boxes_t* b, int t, int ml, double g, const double* r, double* o)
{
    return b_precond<float>(b, t, ml, g, r, o);
}
int or_b_precond_f64(const or_boxes_t* b, int t, int ml, double g, const double* r, double* o)
{
    return b_precond<double>(b, t, ml, g, r, o);
}

long or_b_pairs(const or_boxes_t* b, long cap, int* pi, int* pj)
{
    long cnt = 0;
    for (int i = 0; i < b->nbox; i++)
        for (int j = i; j < b->nbox; j++)
        {
            int s1[3], s2[3];
            if (!ext_shift(b, i, j, s1) && !ext_shift(b, j, i, s2)) continue;
            if (cnt < cap)
            {
                pi[cnt] = i;
                pj[cnt] = j;
            }
            cnt++;
        }
    return cnt;
}

// S_ij = h^3 <psi_i, psi_j>; H_ij = h^3 <psi_i, H psi_j> (Galerkin, as
// LocGridOrbitals::computeLocalProduct forms it from the unmasked H psi,
// src/LocGridOrbitals.cc:1554-1604), from the stored (masked) H psi on the
// overlap plus the stencil's reach 2 points outside each box; the two
// one-sided evaluations are averaged so that H is exactly symmetric.
int or_b_gram_h(const or_boxes_t* b, const double* psi, const double* hpsi, double* S, double* H)
{
    Boxes B(b, 1);
    const int n    = b->nbox;
    const double vel = b->h[0] * b->h[1] * b->h[2];
    C4 c4(b->h);
    std::memset(S, 0, sizeof(double) * n * n);
    std::memset(H, 0, sizeof(double) * n * n);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++)
        {
            const int ei = b->edge[i], ej = b->edge[j];
            const double* pi  = psi + B.off[0][i];
            const double* pj  = psi + B.off[0][j];
            const double* hi  = hpsi + B.off[0][i];
            const double* hj  = hpsi + B.off[0][j];
            double s = 0., h1 = 0., h2 = 0.;
            Ovl o;
            const bool ov = box_overlap(b, i, j, &o);
            if (ov)
                for (int x = o.a[0]; x < o.b[0]; x++)
                    for (int y = o.a[1]; y < o.b[1]; y++)
                        for (int z = o.a[2]; z < o.b[2]; z++)
                        {
                            long ii = ((long)x * ei + y) * ei + z;
                            long jj = ((long)(x + o.sh[0]) * ej + (y + o.sh[1])) * ej + (z + o.sh[2]);
                            s += pi[ii] * pj[jj];
                            h1 += pi[ii] * hj[jj];
                            h2 += hi[ii] * pj[jj];
                        }
            int sij[3], sji[3];
            const bool ex1 = ext_shift(b, i, j, sij);
            const bool ex2 = ext_shift(b, j, i, sji);
            if (!ov && !ex1 && !ex2) continue;
            if (ex1) h1 += skin_sum(pi, ei, pj, ej, sij, c4);
            if (ex2) h2 += skin_sum(pj, ej, pi, ei, sji, c4);
            S[(long)j * n + i] = S[(long)i * n + j] = s * vel;
            H[(long)j * n + i] = H[(long)i * n + j] = 0.5 * (h1 + h2) * vel;
        }
    return 0;
}

int or_b_residual(const or_boxes_t* b, const double* psi, const double* hpsi, const double* theta, double* res)
{
    Boxes B(b, 1);
    const int n = b->nbox;
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < n; j++)
    {
        const int ej = b->edge[j];
        const long nj = (long)ej * ej * ej;
        double* r = res + B.off[0][j];
        for (long p = 0; p < nj; p++)
            r[p] = 0.;
        for (int k = 0; k < n; k++)
        {
            Ovl o;  // j-local ranges, sh maps j-local -> k-local
            if (!box_overlap(b, j, k, &o)) continue;
            const double t  = theta[(long)j * n + k];
            const int ek    = b->edge[k];
            const double* pk = psi + B.off[0][k];
            for (int x = o.a[0]; x < o.b[0]; x++)
                for (int y = o.a[1]; y < o.b[1]; y++)
                    for (int z = o.a[2]; z < o.b[2]; z++)
                        r[((long)x * ej + y) * ej + z]
                            += pk[((long)(x + o.sh[0]) * ek + (y + o.sh[1])) * ek + (z + o.sh[2])] * t;
        }
        const double* hj = hpsi + B.off[0][j];
        for (long p = 0; p < nj; p++)
            r[p] += -1. * hj[p];
    }
    return 0;
}

int or_b_diag_dots(const or_boxes_t* b, const double* r, double* d)
{
    Boxes B(b, 1);
    const double vel = b->h[0] * b->h[1] * b->h[2];
    for (int j = 0; j < b->nbox; j++)
    {
        const double* rj = r + B.off[0][j];
        long nj          = B.off[0][j + 1] - B.off[0][j];
        double acc       = 0.;
        for (long p = 0; p < nj; p++)
            acc += rj[p] * rj[p];
        d[j] = acc * vel;
    }
    return 0;
}

int or_b_axpy(long n, double a, const double* x, double* y)
{
    axpy(n, a, x, y);
    return 0;
}

}  // extern "C"
