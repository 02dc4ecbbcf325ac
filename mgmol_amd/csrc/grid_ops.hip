// grid_ops.hip -- MGmol grid operators on ghosted blocks (drop-in for the
// pb:: layer) and the packed Hamiltonian / preconditioner entry points.
//
// Arithmetic follows the reference kernels operation by operation (the
// library is built with -ffp-contract=off), so FP64 results agree with the
// reference formulas to the last bit wherever the summation order is fixed.
#include "common.hpp"

namespace mgx
{
namespace
{

constexpr double kInv12 = 1. / 12.;
constexpr double kInv64 = 1. / 64.;

__device__ __forceinline__ int wrapi(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// ------------------------------------------------------------------ halo --
// GridFuncVector::trade_boundaries single-task branches
// (src/pb/GridFuncVector.cc:1544-1622, 585-604, 1052-1070, 1426-1440): every
// ghost point receives the periodic image; along a Dirichlet direction
// GridFunc::setBoundaryValues (src/pb/GridFunc.cc:2189-2340) zeroes the ghosts
// and the first interior layer (index 0, the boundary plane of the
// vertex-centred grid), so a point is zero when a Dirichlet coordinate lies
// outside 1 .. n-1.
template <typename T>
__global__ void k_halo(T* u, GGeom g, int64_t total, int bc0, int bc1, int bc2)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.sizeg;
        int64_t r       = t - f * g.sizeg;
        const int x     = (int)(r / g.incx);
        r -= (int64_t)x * g.incx;
        const int y = (int)(r / g.incy);
        const int z = (int)(r - (int64_t)y * g.incy);
        int c[3]    = { x - g.ng, y - g.ng, z - g.ng };
        const int d[3]  = { g.d0, g.d1, g.d2 };
        const int bc[3] = { bc0, bc1, bc2 };
        bool ghost = false, zero = false;
#pragma unroll
        for (int k = 0; k < 3; k++)
        {
            if (bc[k] != 1 && c[k] == 0) zero = true;
            if (c[k] < 0 || c[k] >= d[k])
            {
                ghost = true;
                if (bc[k] == 1)
                    c[k] = wrapi(c[k], d[k]);
                else
                    zero = true;
            }
        }
        if (!ghost && !zero) continue;
        T* uf = u + f * g.sizeg;
        const int64_t dst = (int64_t)x * g.incx + (int64_t)y * g.incy + z;
        const int64_t src = (int64_t)(c[0] + g.ng) * g.incx + (int64_t)(c[1] + g.ng) * g.incy + (c[2] + g.ng);
        uf[dst]           = zero ? (T)0 : uf[src];
    }
}

// k_halo for grids periodic in every direction: only the ghost points, each
// copied from its wrapped interior image (the same values).  Index space:
// first the rows that are ghosts as a whole (the 2 ng ghost x-planes, then
// the 2 ng ghost y-rows of each interior x-plane), na points, then the 2 ng
// z-ghosts of each interior row, nb points.
template <typename T>
__global__ void k_halo_per(T* u, GGeom g, int64_t na, int64_t nb)
{
    const int ng = g.ng;
    const int ey = g.d1 + 2 * ng, ez = (int)g.incy, w = 2 * ng;
    const int64_t rows = (int64_t)w * ey + (int64_t)g.d0 * w;  // whole ghost rows per function
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < na + nb;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        int64_t f;
        int x, y, z;
        if (t < na)
        {
            const int64_t q = t / ez;
            z               = (int)(t - q * ez);
            f               = q / rows;
            int64_t r       = q - f * rows;
            if (r < (int64_t)w * ey)
            {
                const int p = (int)(r / ey);
                y           = (int)(r - (int64_t)p * ey);
                x           = p < ng ? p : g.d0 + p;
            }
            else
            {
                r -= (int64_t)w * ey;
                const int xi = (int)(r / w), k = (int)(r - (int64_t)xi * w);
                x = ng + xi;
                y = k < ng ? k : g.d1 + k;
            }
        }
        else
        {
            int64_t s   = t - na;
            const int k = (int)(s % w);
            s /= w;
            const int yi = (int)(s % g.d1);
            s /= g.d1;
            const int xi = (int)(s % g.d0);
            f            = s / g.d0;
            x = ng + xi;
            y = ng + yi;
            z = k < ng ? k : g.d2 + k;
        }
        T* uf             = u + f * g.sizeg;
        const int cx      = wrapi(x - ng, g.d0) + ng, cy = wrapi(y - ng, g.d1) + ng, cz = wrapi(z - ng, g.d2) + ng;
        uf[(int64_t)x * g.incx + (int64_t)y * g.incy + z] = uf[(int64_t)cx * g.incx + (int64_t)cy * g.incy + cz];
    }
}

// ---------------------------------------------------------------- stencils --
template <typename T>
__device__ __forceinline__ double lap4_at(const T* a, int64_t i, int64_t X, int64_t Y, const Coef4& c,
    double diag)
{
    // FDkernelDel2_4th body, src/pb/FDkernels.cc:140-156
    return diag * (double)a[i] + c.c1x * ((double)a[i - X] + (double)a[i + X])
           + c.c1y * ((double)a[i - Y] + (double)a[i + Y]) + c.c1z * ((double)a[i - 1] + (double)a[i + 1])
           + c.c2x * ((double)a[i - 2 * X] + (double)a[i + 2 * X])
           + c.c2y * ((double)a[i - 2 * Y] + (double)a[i + 2 * Y])
           + c.c2z * ((double)a[i - 2] + (double)a[i + 2]);
}

template <typename T>
__device__ __forceinline__ double lap2_at(const T* v, int64_t i, int64_t X, int64_t Y, const Coef2& c)
{
    // FDkernelDel2_2nd body, src/pb/FDkernels.cc:66-69 (pair sums in T)
    return c.c0 * v[i] + c.c1x * (T)(v[i - X] + v[i + X]) + c.c1y * (T)(v[i - Y] + v[i + Y])
           + c.c1z * (T)(v[i - 1] + v[i + 1]);
}

template <typename T>
__device__ __forceinline__ double lapM_at(const T* v, int64_t i, int64_t X, int64_t Y, const CoefM& c)
{
    // FDkernelDel2_4th_Mehr body, src/pb/FDkernels.cc:498-513
    const T* p = v + i;
    return c.c0 * (double)p[0] + c.cz * (double)(T)(p[-1] + p[1]) + c.cy * (double)(T)(p[-Y] + p[Y])
           + c.cx * (double)(T)(p[-X] + p[X])
           + c.cxz * (double)(T)(p[-X - 1] + p[-X + 1] + p[X - 1] + p[X + 1])
           + c.cyz * (double)(T)(p[-Y - 1] + p[-Y + 1] + p[Y - 1] + p[Y + 1])
           + c.cxy * (double)(T)(p[-X - Y] + p[X - Y] + p[-X + Y] + p[X + Y]);
}

struct AllCoef
{
    Coef4 c4;
    Coef2 c2;
    CoefM cm;
};

template <typename T, int TYPE>
__global__ void k_lap(const T* in, T* out, GGeom g, int64_t total, AllCoef c)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.size;
        int r           = (int)(t - f * g.size);
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y     = r / g.d2;
        const int z     = r - y * g.d2;
        const int64_t i = f * g.sizeg + (int64_t)(x + g.ng) * g.incx + (int64_t)(y + g.ng) * g.incy + z + g.ng;
        double v;
        if (TYPE == MGX_LAP_4TH)
            v = lap4_at(in, i, g.incx, g.incy, c.c4, c.c4.c0);
        else if (TYPE == MGX_LAP_2ND)
            v = lap2_at(in, i, g.incx, g.incy, c.c2);
        else
            v = lapM_at(in, i, g.incx, g.incy, c.cm);
        out[i] = (T)v;
    }
}

// one Mehrstellen Jacobi sweep out of place on the interior: w = -Lap_M x - f
// (written if w is given), xo = x + mjf w, with the arithmetic of k_lap and
// k_jacobi_update in order (the same interior values); x's ghosts are read as
// stored, xo's and w's are not written
template <typename T>
__global__ void k_jacobi_mehr(const T* __restrict__ x, const T* __restrict__ f, T* __restrict__ xo,
    T* __restrict__ w, GGeom g, int64_t total, AllCoef c, double mjf)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t fn = t / g.size;
        int r            = (int)(t - fn * g.size);
        const int px     = r / (g.d1 * g.d2);
        r -= px * g.d1 * g.d2;
        const int py    = r / g.d2;
        const int pz    = r - py * g.d2;
        const int64_t i = fn * g.sizeg + (int64_t)(px + g.ng) * g.incx + (int64_t)(py + g.ng) * g.incy + pz + g.ng;
        T wv            = (T)lapM_at(x, i, g.incx, g.incy, c.cm);
        wv += static_cast<T>(-1. * static_cast<double>(f[i]));
        if (w) w[i] = wv;
        T xv = x[i];
        xv += static_cast<T>(mjf * static_cast<double>(wv));
        xo[i] = xv;
    }
}

// FDkernelRHS_4th_Mehr1, src/pb/FDkernels.cc:523-584
template <typename T>
__global__ void k_rhs(const T* v, T* rhs, GGeom g, GGeom go, int64_t total)
{
    const int64_t X = g.incx, Y = g.incy;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.size;
        int r           = (int)(t - f * g.size);
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y     = r / g.d2;
        const int z     = r - y * g.d2;
        const T* p      = v + f * g.sizeg + (int64_t)(x + g.ng) * X + (int64_t)(y + g.ng) * Y + z + g.ng;
        const int64_t o = f * go.sizeg + (int64_t)(x + go.ng) * go.incx + (int64_t)(y + go.ng) * go.incy + z + go.ng;
        rhs[o] = (T)(0.5 * (double)p[0] + kInv12 * (double)(T)(p[-X] + p[X] + p[-Y] + p[Y] + p[-1] + p[1]));
    }
}

template <typename T>
__global__ void k_pointwise(const T* a, const double* v, T* out, int64_t sizeg, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (T)(a[t] * v[t % sizeg]);
}

template <typename T>
__global__ void k_axpy(int64_t n, double alpha, const T* x, T* y)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        y[t] += static_cast<T>(alpha * static_cast<double>(x[t]));
}

// STREAM triad a = b + s c over FP32 arrays, the HBM ceiling the bandwidth-
// bound kernels are measured against: 16-byte lanes, U independent loads of
// each operand in flight per lane, one pass (n4 float4 values per array)
template <int U>
__global__ void __launch_bounds__(256) k_triad4(const float4* __restrict__ b, const float4* __restrict__ c,
    float4* __restrict__ a, int64_t n4, float sc)
{
    const int64_t base = (int64_t)blockIdx.x * (256 * U) + threadIdx.x;
    float4 vb[U], vc[U];
#pragma unroll
    for (int u = 0; u < U; u++)
    {
        const int64_t i = base + (int64_t)u * 256;
        if (i < n4)
        {
            vb[u] = b[i];
            vc[u] = c[i];
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
    {
        const int64_t i = base + (int64_t)u * 256;
        if (i < n4)
        {
            float4 r;
            r.x  = vb[u].x + sc * vc[u].x;
            r.y  = vb[u].y + sc * vc[u].y;
            r.z  = vb[u].z + sc * vc[u].z;
            r.w  = vb[u].w + sc * vc[u].w;
            a[i] = r;
        }
    }
}

// second half of GridFuncVector::jacobi over all sizeg points:
// w -= f (MPaxpy -1); x += -jf * w (MPaxpy)
template <typename T>
__global__ void k_jacobi_update(int64_t n, double mjf, const T* f, T* w, T* x)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    {
        T wv = w[t];
        wv += static_cast<T>(-1. * static_cast<double>(f[t]));
        w[t] = wv;
        x[t] += static_cast<T>(mjf * static_cast<double>(wv));
    }
}

// MGkernelRestrict3D, src/pb/MGkernels.cc:204-283
template <typename T>
__global__ void k_restrict(const T* uf, GGeom gf, T* uc, GGeom gc, int64_t total)
{
    const int64_t X = gf.incx, Y = gf.incy;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / gc.size;
        int r           = (int)(t - f * gc.size);
        const int x     = r / (gc.d1 * gc.d2);
        r -= x * gc.d1 * gc.d2;
        const int y = r / gc.d2;
        const int z = r - y * gc.d2;
        const T* u0 = uf + f * gf.sizeg + (int64_t)(2 * x + gf.ng) * X + (int64_t)(2 * y + gf.ng) * Y + gf.ng + 2 * z;
        const T* umx = u0 - X;
        const T* upx = u0 + X;
        const T* umy = u0 - Y;
        const T* upy = u0 + Y;
        const T* umxpy = u0 - X + Y;
        const T* upxpy = u0 + X + Y;
        const T* umxmy = u0 - X - Y;
        const T* upxmy = u0 + X - Y;
        double face = (double)upx[0] + (double)umx[0] + (double)upy[0] + (double)umy[0] + (double)u0[-1]
                      + (double)u0[1];
        double corner = (double)upxpy[-1] + (double)upxpy[1] + (double)upxmy[-1] + (double)upxmy[1]
                        + (double)umxpy[-1] + (double)umxpy[1] + (double)umxmy[-1] + (double)umxmy[1];
        double edge = (double)upy[-1] + (double)upy[1] + (double)umy[-1] + (double)umy[1] + (double)upx[-1]
                      + (double)upx[1] + (double)umx[-1] + (double)umx[1] + (double)umxmy[0] + (double)upxmy[0]
                      + (double)umxpy[0] + (double)upxpy[0];
        uc[f * gc.sizeg + (int64_t)(x + gc.ng) * gc.incx + (int64_t)(y + gc.ng) * gc.incy + z + gc.ng]
            = (T)(kInv64 * (8. * u0[0] + 4. * face + 2. * edge + corner));
    }
}

// MGkernelExtend3D injection, src/pb/MGkernels.cc:59-82 (includes the first
// ghost layer on the + side)
template <typename T>
__global__ void k_inject(const T* uc, GGeom gc, T* uf, GGeom gf, int64_t total)
{
    const int a0 = gc.d0 + 1, a1 = gc.d1 + 1, a2 = gc.d2 + 1;
    const int64_t per = (int64_t)a0 * a1 * a2;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / per;
        int r           = (int)(t - f * per);
        const int x     = r / (a1 * a2);
        r -= x * a1 * a2;
        const int y = r / a2;
        const int z = r - y * a2;
        uf[f * gf.sizeg + (int64_t)(2 * x + gf.ng) * gf.incx + (int64_t)(2 * y + gf.ng) * gf.incy + 2 * z + gf.ng]
            = uc[f * gc.sizeg + (int64_t)(x + gc.ng) * gc.incx + (int64_t)(y + gc.ng) * gc.incy + z + gc.ng];
    }
}

// MGkernelExtend3D interpolation, src/pb/MGkernels.cc:85-195: every odd
// point reads only injected (even,even,even) points, in T arithmetic.
template <typename T>
__device__ __forceinline__ T prolong_point(const T* fd, int64_t i, int ox, int oy, int oz, int64_t X, int64_t Y)
{
    if (!ox && !oy && !oz) return fd[i];
    if (!ox && !oy && oz) return (T)(0.5 * (T)(fd[i - 1] + fd[i + 1]));
    if (!ox && oy && !oz) return (T)(0.5 * (T)(fd[i + Y] + fd[i - Y]));
    if (!ox && oy && oz) return (T)(0.25 * (T)(fd[i + 1 + Y] + fd[i + 1 - Y] + fd[i - 1 + Y] + fd[i - 1 - Y]));
    if (ox && !oy && oz) return (T)(0.25 * (T)(fd[i + X + 1] + fd[i + X - 1] + fd[i - X + 1] + fd[i - X - 1]));
    if (ox && !oy && !oz) return (T)(0.5 * (T)(fd[i + X] + fd[i - X]));
    if (ox && oy && oz)
        return (T)(0.125
                   * (T)(fd[i + X + Y + 1] + fd[i + X + Y - 1] + fd[i + X - Y + 1] + fd[i + X - Y - 1]
                         + fd[i - X + Y + 1] + fd[i - X + Y - 1] + fd[i - X - Y + 1] + fd[i - X - Y - 1]));
    return (T)(0.25 * (T)(fd[i + X + Y] + fd[i + X - Y] + fd[i - X + Y] + fd[i - X - Y]));
}

template <typename T>
__global__ void k_fill_odd(T* uf, GGeom gf, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / gf.size;
        int r           = (int)(t - f * gf.size);
        const int x     = r / (gf.d1 * gf.d2);
        r -= x * gf.d1 * gf.d2;
        const int y = r / gf.d2;
        const int z = r - y * gf.d2;
        const int ox = x & 1, oy = y & 1, oz = z & 1;
        if (!ox && !oy && !oz) continue;
        T* fd           = uf + f * gf.sizeg;
        const int64_t i = (int64_t)(x + gf.ng) * gf.incx + (int64_t)(y + gf.ng) * gf.incy + z + gf.ng;
        fd[i]           = prolong_point(fd, i, ox, oy, oz, gf.incx, gf.incy);
    }
}

// The coarse correction of a V-cycle in one pass: x -= mask * P(uc) on the
// fine interior, P the MGkernelExtend3D interpolation (src/pb/MGkernels.cc:
// 85-195) formed from the coarse values directly (the fine even points are
// the injected coarse points, so every term and its order is that of
// k_inject + k_fill_odd), the subtraction as k_axpy with alpha = -1.  One
// block row per fine (function, x) plane; uc must have fresh ghosts.
template <typename T>
__global__ void __launch_bounds__(256) k_prolong_sub(const T* __restrict__ uc, GGeom gc, T* __restrict__ xf,
    GGeom gf, const uint8_t* __restrict__ mask)
{
    const int plane = blockIdx.y;  // f * d0 + x
    const int f = plane / gf.d0, x = plane - (plane / gf.d0) * gf.d0;
    const int np = gf.d1 * gf.d2;
    const int p  = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int y = p / gf.d2, z = p - (p / gf.d2) * gf.d2;
    const int64_t t = ((int64_t)f * gf.d0 + x) * np + p;
    if (mask && !mask[t]) return;
    const int ox = x & 1, oy = y & 1, oz = z & 1;
    const T* c = uc + (int64_t)f * gc.sizeg + (int64_t)((x >> 1) + gc.ng) * gc.incx
                 + (int64_t)((y >> 1) + gc.ng) * gc.incy + (z >> 1) + gc.ng;
    // fine +-1 along an odd direction: coarse +1 / +0
    const int64_t X = gc.incx, Y = gc.incy;
    T v;
    if (!ox && !oy && !oz)
        v = c[0];
    else if (!ox && !oy && oz)
        v = (T)(0.5 * (T)(c[0] + c[1]));
    else if (!ox && oy && !oz)
        v = (T)(0.5 * (T)(c[Y] + c[0]));
    else if (!ox && oy && oz)
        v = (T)(0.25 * (T)(c[1 + Y] + c[1] + c[Y] + c[0]));
    else if (ox && !oy && oz)
        v = (T)(0.25 * (T)(c[X + 1] + c[X] + c[1] + c[0]));
    else if (ox && !oy && !oz)
        v = (T)(0.5 * (T)(c[X] + c[0]));
    else if (ox && oy && oz)
        v = (T)(0.125 * (T)(c[X + Y + 1] + c[X + Y] + c[X + 1] + c[X] + c[Y + 1] + c[Y] + c[1] + c[0]));
    else
        v = (T)(0.25 * (T)(c[X + Y] + c[X] + c[Y] + c[0]));
    T* xp = xf + (int64_t)f * gf.sizeg + (int64_t)(x + gf.ng) * gf.incx + (int64_t)(y + gf.ng) * gf.incy + z + gf.ng;
    *xp += static_cast<T>(-1. * static_cast<double>(v));
}

template <typename T>
__global__ void k_mask(T* u, const uint8_t* m, GGeom g, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        if (m[t]) continue;
        const int64_t f = t / g.size;
        int r           = (int)(t - f * g.size);
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y = r / g.d2;
        const int z = r - y * g.d2;
        u[f * g.sizeg + (int64_t)(x + g.ng) * g.incx + (int64_t)(y + g.ng) * g.incy + z + g.ng] = (T)0;
    }
}

template <typename T>
__global__ void k_copy_interior(const T* src, T* dst, GGeom g, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.size;
        int r           = (int)(t - f * g.size);
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y     = r / g.d2;
        const int z     = r - y * g.d2;
        const int64_t i = f * g.sizeg + (int64_t)(x + g.ng) * g.incx + (int64_t)(y + g.ng) * g.incy + z + g.ng;
        dst[i]          = src[i];
    }
}

template <typename T>
__global__ void k_fill(T* u, int64_t n, T v)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        u[t] = v;
}

// packed FP64 -> ghosted T (zero ghosts): BlockVector::setDataWithGhosts,
// src/BlockVector.cc:489-517 (resetData then assign)
template <typename T>
__global__ void k_packed_to_ghosted(const double* src, int64_t ld, T* dst, GGeom g, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.sizeg;
        int64_t r       = t - f * g.sizeg;
        const int x     = (int)(r / g.incx) - g.ng;
        r -= (int64_t)(x + g.ng) * g.incx;
        const int y = (int)(r / g.incy) - g.ng;
        const int z = (int)(r - (int64_t)(y + g.ng) * g.incy) - g.ng;
        T v         = (T)0;
        if (x >= 0 && y >= 0 && z >= 0 && x < g.d0 && y < g.d1 && z < g.d2)
            v = (T)src[f * ld + ((int64_t)x * g.d1 + y) * g.d2 + z];
        dst[t] = v;
    }
}

// k_packed_to_ghosted into f with the preconditioner's start x = 0 + gamma f
// (the fill and k_axpy it replaces, the same arithmetic) in the same pass
template <typename T>
__global__ void k_packed_to_ghosted_start(const double* src, int64_t ld, T* dst, T* xs, double gamma, GGeom g,
    int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.sizeg;
        int64_t r       = t - f * g.sizeg;
        const int x     = (int)(r / g.incx) - g.ng;
        r -= (int64_t)(x + g.ng) * g.incx;
        const int y = (int)(r / g.incy) - g.ng;
        const int z = (int)(r - (int64_t)(y + g.ng) * g.incy) - g.ng;
        T v         = (T)0;
        if (x >= 0 && y >= 0 && z >= 0 && x < g.d0 && y < g.d1 && z < g.d2)
            v = (T)src[f * ld + ((int64_t)x * g.d1 + y) * g.d2 + z];
        dst[t] = v;
        T xv   = (T)0;
        xv += static_cast<T>(gamma * static_cast<double>(v));
        xs[t] = xv;
    }
}

template <typename T>
__global__ void k_ghosted_to_packed(const T* src, GGeom g, double* dst, int64_t ld, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.size;
        int r           = (int)(t - f * g.size);
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y = r / g.d2;
        const int z = r - y * g.d2;
        dst[f * ld + ((int64_t)x * g.d1 + y) * g.d2 + z]
            = (double)src[f * g.sizeg + (int64_t)(x + g.ng) * g.incx + (int64_t)(y + g.ng) * g.incy + z + g.ng];
    }
}

// --------------------------------------------- packed Hamiltonian apply ---
struct PGeom
{
    int d0, d1, d2;
    int bc0, bc1, bc2;
    int64_t size;
};

// the neighbour c + d along one direction, false when its value is zero: a
// Dirichlet direction (bc 0) holds zero outside 1 .. n-1 (the index-0 layer
// is the boundary plane, as after a halo exchange)
__device__ __forceinline__ bool nb(int c, int d, int n, int bc, int* out)
{
    int v = c + d;
    if (!bc && v == 0) return false;
    if (v < 0 || v >= n)
    {
        if (!bc) return false;
        v = wrapi(v, n);
    }
    *out = v;
    return true;
}

template <typename T>
__device__ __forceinline__ double pval(const T* p, const PGeom& g, int x, int y, int z, int dx, int dy, int dz)
{
    int a, b, c;
    if (!nb(x, dx, g.d0, g.bc0, &a) || !nb(y, dy, g.d1, g.bc1, &b) || !nb(z, dz, g.d2, g.bc2, &c)) return 0.;
    return (double)p[((int64_t)a * g.d1 + b) * g.d2 + c];
}

template <typename T>
__device__ __forceinline__ T pvalT(const T* p, const PGeom& g, int x, int y, int z, int dx, int dy, int dz)
{
    int a, b, c;
    if (!nb(x, dx, g.d0, g.bc0, &a) || !nb(y, dy, g.d1, g.bc1, &b) || !nb(z, dz, g.d2, g.bc2, &c)) return (T)0;
    return p[((int64_t)a * g.d1 + b) * g.d2 + c];
}

// FDoper<T>::del2_4th_withPot, src/pb/FDoper.cc:322-399, on packed data with
// the halo applied by index wrapping.
template <typename T>
__global__ void k_hloc4_packed(const T* psi, int64_t ld, const double* vtot, T* hpsi, int64_t ldo, PGeom g,
    Coef4 c, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.size;
        const int64_t p = t - f * g.size;
        int r           = (int)p;
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y  = r / g.d2;
        const int z  = r - y * g.d2;
        const T* a   = psi + f * ld;
        double v     = (c.c0 + vtot[p]) * pval(a, g, x, y, z, 0, 0, 0)
                   + c.c1x * (pval(a, g, x, y, z, -1, 0, 0) + pval(a, g, x, y, z, 1, 0, 0))
                   + c.c1y * (pval(a, g, x, y, z, 0, -1, 0) + pval(a, g, x, y, z, 0, 1, 0))
                   + c.c1z * (pval(a, g, x, y, z, 0, 0, -1) + pval(a, g, x, y, z, 0, 0, 1))
                   + c.c2x * (pval(a, g, x, y, z, -2, 0, 0) + pval(a, g, x, y, z, 2, 0, 0))
                   + c.c2y * (pval(a, g, x, y, z, 0, -2, 0) + pval(a, g, x, y, z, 0, 2, 0))
                   + c.c2z * (pval(a, g, x, y, z, 0, 0, -2) + pval(a, g, x, y, z, 0, 0, 2));
        hpsi[f * ldo + p] = (T)v;
    }
}

// Hamiltonian<T>::applyLocal Mehrstellen branch, src/Hamiltonian.cc:106-131:
// w1 = psi*V (pointwise), work1 = B w1, w1 = L_M psi, work1 += w1.
template <typename T>
__device__ __forceinline__ T vpsi(const T* a, const double* v, const PGeom& g, int x, int y, int z, int dx, int dy,
    int dz)
{
    int aa, b, c;
    if (!nb(x, dx, g.d0, g.bc0, &aa) || !nb(y, dy, g.d1, g.bc1, &b) || !nb(z, dz, g.d2, g.bc2, &c)) return (T)0;
    const int64_t q = ((int64_t)aa * g.d1 + b) * g.d2 + c;
    return (T)(a[q] * v[q]);
}

template <typename T>
__global__ void k_hlocM_packed(const T* psi, int64_t ld, const double* vtot, T* hpsi, int64_t ldo, PGeom g,
    CoefM c, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.size;
        const int64_t p = t - f * g.size;
        int r           = (int)p;
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y = r / g.d2;
        const int z = r - y * g.d2;
        const T* a  = psi + f * ld;
        const T bw  = (T)(0.5 * (double)vpsi(a, vtot, g, x, y, z, 0, 0, 0)
                         + kInv12
                               * (double)(T)(vpsi(a, vtot, g, x, y, z, -1, 0, 0) + vpsi(a, vtot, g, x, y, z, 1, 0, 0)
                                             + vpsi(a, vtot, g, x, y, z, 0, -1, 0) + vpsi(a, vtot, g, x, y, z, 0, 1, 0)
                                             + vpsi(a, vtot, g, x, y, z, 0, 0, -1)
                                             + vpsi(a, vtot, g, x, y, z, 0, 0, 1)));
#define P(dx, dy, dz) pvalT(a, g, x, y, z, dx, dy, dz)
        const T lm = (T)(c.c0 * (double)P(0, 0, 0) + c.cz * (double)(T)(P(0, 0, -1) + P(0, 0, 1))
                         + c.cy * (double)(T)(P(0, -1, 0) + P(0, 1, 0)) + c.cx * (double)(T)(P(-1, 0, 0) + P(1, 0, 0))
                         + c.cxz * (double)(T)(P(-1, 0, -1) + P(-1, 0, 1) + P(1, 0, -1) + P(1, 0, 1))
                         + c.cyz * (double)(T)(P(0, -1, -1) + P(0, -1, 1) + P(0, 1, -1) + P(0, 1, 1))
                         + c.cxy * (double)(T)(P(-1, -1, 0) + P(1, -1, 0) + P(-1, 1, 0) + P(1, 1, 0)));
#undef P
        T out = bw;
        out += static_cast<T>(1. * static_cast<double>(lm));
        hpsi[f * ldo + p] = out;
    }
}

// FDkernelRHS_4th_Mehr1 (src/pb/FDkernels.cc:523-584), i.e. Lap::rhs of the
// Mehrstellen operator, on packed data with the halo by index wrapping:
// out = (T)(0.5 psi + (1/12) (T)(sum of the 6 face neighbours))
template <typename T>
__global__ void k_rhsM_packed(const T* psi, int64_t ld, T* out, int64_t ldo, PGeom g, int64_t total)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / g.size;
        const int64_t p = t - f * g.size;
        int r           = (int)p;
        const int x     = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y = r / g.d2;
        const int z = r - y * g.d2;
        const T* a  = psi + f * ld;
#define P(dx, dy, dz) pvalT(a, g, x, y, z, dx, dy, dz)
        out[f * ldo + p] = (T)(0.5 * (double)P(0, 0, 0)
                               + kInv12 * (double)(T)(P(-1, 0, 0) + P(1, 0, 0) + P(0, -1, 0) + P(0, 1, 0)
                                                      + P(0, 0, -1) + P(0, 0, 1)));
#undef P
    }
}

// ------------------------------------------------------ marching sweeps --
// The 4th- and 2nd-order stencils on whole grids, marching along x: a block
// owns a tile of TY y-rows (all z) of one function over a range of x-planes
// (an x-segment).  Each lane owns pairs of z-points (8 or 16 bytes) and keeps
// them for the 2R + 1 planes x - R .. x + R in registers; the centre plane of
// the tile, plus R halo rows on each y side and R z-ghosts on each row end,
// goes through LDS (two buffers, one barrier per plane).  WRAP: neighbours
// outside the grid are periodic images (bc = 1) or zero (bc = 0) and the
// input's ghosts are never read (the sweep needs no halo exchange before it);
// otherwise the input's ghost cells are read as stored.  Arithmetic per
// point is that of lap4_at / lap2_at / k_hloc4_packed / k_jacobi_update, in
// the same order.
struct GMarch
{
    int d0, d1, d2;
    int bc0, bc1, bc2;
    int64_t ix, iy, i0, istr;  // input (and f): strides, offset of interior (0,0,0), function stride
    int64_t ox, oy, o0, ostr;  // output (and w)
    int ty, nty, nxs, xlen;    // tile rows, y tiles, x segments, planes per segment
    int64_t nreal;             // blocks with work (the grid is padded to a multiple of 8)
};

constexpr int MG_LAP = 0, MG_JACOBI = 1, MG_HLOC = 2;

template <typename T, int TYPE, int MODE, bool WRAP, int NT, int PM = 1, int PD = 1>
__global__ void __launch_bounds__(NT) k_grid_march(const T* __restrict__ in, const T* __restrict__ f,
    const double* __restrict__ vtot, T* __restrict__ out, T* __restrict__ wout, GMarch g, Coef4 c4, Coef2 c2,
    double mjf)
{
    typedef T T2 __attribute__((ext_vector_type(2)));
    constexpr int R  = TYPE == MGX_LAP_4TH ? 2 : 1;  // stencil reach
    constexpr int NQ = 2 * R + 1;
    extern __shared__ unsigned char lds_raw[];
    T* lds0 = reinterpret_cast<T*>(lds_raw);
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    if (L >= g.nreal) return;
    const int per_f = g.nty * g.nxs;
    const int fn    = (int)(L / per_f);
    const int rem   = (int)(L - (int64_t)fn * per_f);
    const int seg = rem / g.nty, tyi = rem - (rem / g.nty) * g.nty;
    const int y0 = tyi * g.ty, ny = min(g.ty, g.d1 - y0);
    const int xs = seg * g.xlen, xe = min(xs + g.xlen, g.d0);
    const int G   = g.d2 >> 1;
    const int LW  = g.d2 + 4;            // LDS row: z = -2 .. d2 + 1
    const int PLS = (g.ty + 2 * R) * LW;  // one plane buffer
    const T* inf = in + (int64_t)fn * g.istr + g.i0;
    const T* fff = f ? f + (int64_t)fn * g.istr + g.i0 : nullptr;
    T* of        = out + (int64_t)fn * g.ostr + g.o0;
    T* wf        = wout ? wout + (int64_t)fn * g.ostr + g.o0 : nullptr;
    T2 zero2;
    zero2[0] = (T)0;
    zero2[1] = (T)0;
    // x / y coordinates of a neighbour (wrapped), false when it is zero (a
    // Dirichlet direction holds zero outside 1 .. n-1)
    auto cx = [&](int x, int& r) -> bool {
        if (!WRAP || (x > 0 && x < g.d0))
        {
            r = x;
            return true;
        }
        if (!g.bc0) return false;
        r = x < 0 ? x + g.d0 : (x >= g.d0 ? x - g.d0 : x);
        return true;
    };
    auto cy = [&](int y, int& r) -> bool {
        if (!WRAP || (y > 0 && y < g.d1))
        {
            r = y;
            return true;
        }
        if (!g.bc1) return false;
        r = y < 0 ? y + g.d1 : (y >= g.d1 ? y - g.d1 : y);
        return true;
    };
    auto ld2 = [&](int x, int y, int z) -> T2 {
        int xx, yy;
        if (!cx(x, xx) || !cy(y, yy)) return zero2;
        T2 v = *reinterpret_cast<const T2*>(inf + (int64_t)xx * g.ix + (int64_t)yy * g.iy + z);
        if (WRAP && !g.bc2 && z == 0) v[0] = (T)0;
        return v;
    };
    // the z-ghost pair of a row end (side 0: z = -2, -1; side 1: z = d2, d2 + 1)
    auto ghost2 = [&](int x, int y, int side, const T2& own) -> T2 {
        if (WRAP) return g.bc2 ? own : zero2;
        int xx, yy;
        cx(x, xx);
        cy(y, yy);
        const T* p = inf + (int64_t)xx * g.ix + (int64_t)yy * g.iy + (side ? g.d2 : -2);
        T2 v;
        v[0] = p[0];
        v[1] = p[1];
        return v;
    };
    int qy[PM], qz[PM];
    bool act[PM];
#pragma unroll
    for (int k = 0; k < PM; k++)
    {
        const int item = threadIdx.x + k * NT;
        act[k]         = item < ny * G;
        qy[k]          = act[k] ? item / G : 0;
        qz[k]          = act[k] ? 2 * (item - (item / G) * G) : 0;
    }
    // halo rows: 2R rows of G pairs, at most 2 per thread
    const int nh = 2 * R * G;
    int hy[2], hz[2];
    bool hact[2];
#pragma unroll
    for (int k = 0; k < 2; k++)
    {
        const int item = threadIdx.x + k * NT;
        hact[k]        = item < nh;
        const int hr   = hact[k] ? item / G : 0;  // 0 .. 2R-1
        hz[k]          = hact[k] ? 2 * (item - (item / G) * G) : 0;
        hy[k]          = hr < R ? y0 - R + hr : y0 + ny + (hr - R);
    }
    // zero the ghost columns once (Dirichlet z under WRAP never writes them)
    for (int t = threadIdx.x; t < 2 * PLS; t += NT)
        lds0[t] = (T)0;
    T2 q[PM][NQ];
#pragma unroll
    for (int k = 0; k < PM; k++)
#pragma unroll
        for (int j = 0; j < NQ; j++)
            q[k][j] = act[k] ? ld2(xs - R + j, y0 + qy[k], qz[k]) : zero2;
    // loads PD planes ahead: pf[i] = own pairs of plane x + R + 1 + i, hf[i]
    // = halo pairs of plane x + 1 + i
    T2 pf[PD][PM], hv[2], hf[PD][2];
#pragma unroll
    for (int i = 0; i < PD; i++)
    {
#pragma unroll
        for (int k = 0; k < PM; k++)
            pf[i][k] = (act[k] && xs + 1 + i < xe) ? ld2(xs + R + 1 + i, y0 + qy[k], qz[k]) : zero2;
#pragma unroll
        for (int k = 0; k < 2; k++)
            hf[i][k] = (hact[k] && xs + 1 + i < xe) ? ld2(xs + 1 + i, hy[k], hz[k]) : zero2;
    }
#pragma unroll
    for (int k = 0; k < 2; k++)
        hv[k] = hact[k] ? ld2(xs, hy[k], hz[k]) : zero2;
    __syncthreads();
    for (int x = xs; x < xe; x++)
    {
        T* pl = lds0 + ((x - xs) & 1) * PLS;
        // the centre plane into LDS: own pairs, halo pairs, row-end ghosts
#pragma unroll
        for (int k = 0; k < PM; k++)
        {
            if (!act[k]) continue;
            T* row = pl + (qy[k] + R) * LW + 2;
            *reinterpret_cast<T2*>(row + qz[k]) = q[k][R];
            if (qz[k] == 0) *reinterpret_cast<T2*>(row + g.d2) = ghost2(x, y0 + qy[k], 1, q[k][R]);
            if (qz[k] == g.d2 - 2) *reinterpret_cast<T2*>(row - 2) = ghost2(x, y0 + qy[k], 0, q[k][R]);
        }
#pragma unroll
        for (int k = 0; k < 2; k++)
        {
            if (!hact[k]) continue;
            const int hr = hy[k] - (y0 - R);
            T* row       = pl + hr * LW + 2;
            *reinterpret_cast<T2*>(row + hz[k]) = hv[k];
            if (hz[k] == 0) *reinterpret_cast<T2*>(row + g.d2) = ghost2(x, hy[k], 1, hv[k]);
            if (hz[k] == g.d2 - 2) *reinterpret_cast<T2*>(row - 2) = ghost2(x, hy[k], 0, hv[k]);
        }
        __syncthreads();
        // the planes the step after next-PD needs: own pairs at x + R + 1 +
        // PD, halo pairs at x + 1 + PD
        T2 nxt[PM], hnx[2];
#pragma unroll
        for (int k = 0; k < PM; k++)
            nxt[k] = (act[k] && x + 1 + PD < xe) ? ld2(x + R + 1 + PD, y0 + qy[k], qz[k]) : zero2;
#pragma unroll
        for (int k = 0; k < 2; k++)
            hnx[k] = (hact[k] && x + 1 + PD < xe) ? ld2(x + 1 + PD, hy[k], hz[k]) : zero2;
#pragma unroll
        for (int k = 0; k < PM; k++)
        {
            if (!act[k]) continue;
            const int y = y0 + qy[k], z = qz[k];
            const T* ctr = pl + (qy[k] + R) * LW + 2 + z;
            const T2 cc  = q[k][R];
            const T2 ym1 = *reinterpret_cast<const T2*>(ctr - LW);
            const T2 yp1 = *reinterpret_cast<const T2*>(ctr + LW);
            const T2 zl  = *reinterpret_cast<const T2*>(ctr - 2);
            const T2 zr  = *reinterpret_cast<const T2*>(ctr + 2);
            T2 ym2 = zero2, yp2 = zero2;
            if (TYPE == MGX_LAP_4TH)
            {
                ym2 = *reinterpret_cast<const T2*>(ctr - 2 * LW);
                yp2 = *reinterpret_cast<const T2*>(ctr + 2 * LW);
            }
            // z-line around the pair: zz[0..5] = z-2 .. z+3
            const T zz[6] = { zl[0], zl[1], cc[0], cc[1], zr[0], zr[1] };
            T2 fv = zero2;
            if (MODE == MG_JACOBI) fv = *reinterpret_cast<const T2*>(fff + (int64_t)x * g.ix + (int64_t)y * g.iy + z);
            typedef double D2 __attribute__((ext_vector_type(2)));
            D2 vv = { 0., 0. };
            if (MODE == MG_HLOC) vv = *reinterpret_cast<const D2*>(vtot + ((int64_t)x * g.d1 + y) * g.d2 + z);
            T2 ov, wv2;
#pragma unroll
            for (int i = 0; i < 2; i++)
            {
                double r;
                if (TYPE == MGX_LAP_4TH)
                {
                    double diag = c4.c0;
                    if (MODE == MG_HLOC) diag = c4.c0 + vv[i];
                    r = diag * (double)cc[i] + c4.c1x * ((double)q[k][1][i] + (double)q[k][3][i])
                        + c4.c1y * ((double)ym1[i] + (double)yp1[i])
                        + c4.c1z * ((double)zz[i + 1] + (double)zz[i + 3])
                        + c4.c2x * ((double)q[k][0][i] + (double)q[k][4 % NQ][i])
                        + c4.c2y * ((double)ym2[i] + (double)yp2[i]) + c4.c2z * ((double)zz[i] + (double)zz[i + 4]);
                }
                else
                {
                    r = c2.c0 * cc[i] + c2.c1x * (T)(q[k][0][i] + q[k][2][i]) + c2.c1y * (T)(ym1[i] + yp1[i])
                        + c2.c1z * (T)(zz[i + 1] + zz[i + 3]);
                }
                if (MODE == MG_JACOBI)
                {
                    T wv = (T)r;
                    wv += static_cast<T>(-1. * static_cast<double>(fv[i]));
                    wv2[i] = wv;
                    T xv   = cc[i];
                    xv += static_cast<T>(mjf * static_cast<double>(wv));
                    ov[i] = xv;
                }
                else
                    ov[i] = (T)r;
            }
            const int64_t o = (int64_t)x * g.ox + (int64_t)y * g.oy + z;
            *reinterpret_cast<T2*>(of + o) = ov;
            if (MODE == MG_JACOBI && wf) *reinterpret_cast<T2*>(wf + o) = wv2;
        }
#pragma unroll
        for (int k = 0; k < PM; k++)
        {
#pragma unroll
            for (int j = 0; j + 1 < NQ; j++)
                q[k][j] = q[k][j + 1];
            q[k][NQ - 1] = pf[0][k];
#pragma unroll
            for (int i = 0; i + 1 < PD; i++)
                pf[i][k] = pf[i + 1][k];
            pf[PD - 1][k] = nxt[k];
        }
#pragma unroll
        for (int k = 0; k < 2; k++)
        {
            hv[k] = hf[0][k];
#pragma unroll
            for (int i = 0; i + 1 < PD; i++)
                hf[i][k] = hf[i + 1][k];
            hf[PD - 1][k] = hnx[k];
        }
    }
}

// --------------------------------------------------------------- drivers --
// boundary conditions per direction: 1 periodic, 0 Dirichlet
int check_bc(mgx_ctx* ctx, const mgx_grid_t& g)
{
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, g.bc[i] == 0 || g.bc[i] == 1,
            "boundary conditions are 0 (Dirichlet) or 1 (periodic); multipole values are not supported");
    return MGX_OK;
}

int check_block(mgx_ctx* ctx, const mgx_block_t* b)
{
    MGX_REQUIRE(ctx, b && b->data, "null block");
    MGX_REQUIRE(ctx, b->dtype == MGX_F32 || b->dtype == MGX_F64, "bad dtype");
    MGX_REQUIRE(ctx, b->nfunc > 0, "nfunc must be positive");
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, b->grid.dim[i] > 0 && b->grid.dim[i] < 100000, "bad grid dim");
    MGX_REQUIRE(ctx, b->grid.ghosts >= 0, "bad ghost width");
    return check_bc(ctx, b->grid);
}

bool same_shape(const mgx_block_t* a, const mgx_block_t* b)
{
    return a->dtype == b->dtype && a->nfunc == b->nfunc && a->grid.ghosts == b->grid.ghosts
           && a->grid.dim[0] == b->grid.dim[0] && a->grid.dim[1] == b->grid.dim[1]
           && a->grid.dim[2] == b->grid.dim[2];
}

template <typename T>
int launch_halo(mgx_ctx* ctx, T* u, int nfunc, const mgx_grid_t& gr)
{
    GGeom g = ggeom(gr);
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, gr.ghosts <= gr.dim[i], "ghost width exceeds dimension");
    if (gr.bc[0] == 1 && gr.bc[1] == 1 && gr.bc[2] == 1)
    {
        if (g.ng == 0) return MGX_OK;
        const int w        = 2 * g.ng;
        const int64_t rows = (int64_t)w * (g.d1 + w) + (int64_t)g.d0 * w;
        const int64_t na = (int64_t)nfunc * rows * g.incy, nb = (int64_t)nfunc * g.d0 * g.d1 * w;
        hipLaunchKernelGGL(k_halo_per<T>, dim3(nblocks(na + nb)), dim3(kBlock), 0, ctx->stream, u, g, na, nb);
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    const int64_t total = (int64_t)nfunc * g.sizeg;
    hipLaunchKernelGGL(k_halo<T>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, u, g, total, gr.bc[0],
        gr.bc[1], gr.bc[2]);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// geometry of a block for the marching sweeps: strides, offset of interior
// point (0,0,0), function stride
struct MView
{
    int64_t ix, iy, i0, str;
};

inline MView mview_ghosted(const mgx_grid_t& gr)
{
    const GGeom g = ggeom(gr);
    return MView{ g.incx, g.incy, (int64_t)g.ng * (g.incx + g.incy + 1), g.sizeg };
}

inline MView mview_packed(const mgx_grid_t& gr, int64_t ld)
{
    return MView{ (int64_t)gr.dim[1] * gr.dim[2], gr.dim[2], 0, ld };
}

template <typename T, int TYPE, int MODE, bool WRAP>
void march_grid_nt(mgx_ctx* ctx, int items, dim3 grid, size_t lds, const T* in, const T* f, const double* v, T* out,
    T* w, const GMarch& g, const Coef4& c4, const Coef2& c2, double mjf)
{
    // 256 threads, each with up to four lane groups of the tile's plane
    const int pm = (items + 255) / 256;
    if (pm <= 1)
        hipLaunchKernelGGL((k_grid_march<T, TYPE, MODE, WRAP, 256, 1>), grid, dim3(256), lds, ctx->stream, in, f, v,
            out, w, g, c4, c2, mjf);
    else if (pm == 2)
        hipLaunchKernelGGL((k_grid_march<T, TYPE, MODE, WRAP, 256, 2>), grid, dim3(256), lds, ctx->stream, in, f, v,
            out, w, g, c4, c2, mjf);
    else if (pm == 3)
        hipLaunchKernelGGL((k_grid_march<T, TYPE, MODE, WRAP, 256, 3>), grid, dim3(256), lds, ctx->stream, in, f, v,
            out, w, g, c4, c2, mjf);
    else
        hipLaunchKernelGGL((k_grid_march<T, TYPE, MODE, WRAP, 256, 4>), grid, dim3(256), lds, ctx->stream, in, f, v,
            out, w, g, c4, c2, mjf);
}

// The marching sweep (k_grid_march) when it applies: 4th- or 2nd-order
// operator, even z extent up to 256, even strides and offsets, 2-element
// aligned pointers; MG_LAP reads the input's ghosts (at least 2), MG_JACOBI
// and MG_HLOC wrap by the boundary conditions.  Returns 1 if launched, 0 if
// not applicable (the caller then uses the point kernels), or an error.
template <typename T>
int march_grid(mgx_ctx* ctx, int type, int mode, const T* in, const MView& vi, const T* f, const double* vtot, T* out,
    const MView& vo, T* w, int nfunc, const mgx_grid_t& gr, double mjf)
{
    if (type != MGX_LAP_4TH && type != MGX_LAP_2ND) return 0;
    if (mode == MG_HLOC && type != MGX_LAP_4TH) return 0;
    const bool wrap = mode != MG_LAP;
    const int d0 = gr.dim[0], d1 = gr.dim[1], d2 = gr.dim[2];
    if (d2 % 2 || d2 > 256 || d0 < 3 || d1 < 2) return 0;
    if (!wrap && gr.ghosts < 2) return 0;
    for (int64_t v : { vi.ix, vi.iy, vi.i0, vi.str, vo.ix, vo.iy, vo.i0, vo.str })
        if (v & 1) return 0;
    const uintptr_t al = 2 * sizeof(T) - 1;
    for (const void* p : { (const void*)in, (const void*)out, (const void*)f, (const void*)w })
        if (reinterpret_cast<uintptr_t>(p) & al) return 0;
    const int G = d2 / 2, R = type == MGX_LAP_4TH ? 2 : 1;
    GMarch g;
    g.d0 = d0;
    g.d1 = d1;
    g.d2 = d2;
    g.bc0 = gr.bc[0] == 1;
    g.bc1 = gr.bc[1] == 1;
    g.bc2 = gr.bc[2] == 1;
    g.ix = vi.ix;
    g.iy = vi.iy;
    g.i0 = vi.i0;
    g.istr = vi.str;
    g.ox = vo.ix;
    g.oy = vo.iy;
    g.o0 = vo.i0;
    g.ostr = vo.str;
    // tile rows: at most 16, and few enough z-pairs per block that each of
    // the 256 threads holds one lane group (H psi) or two (the sweeps and
    // -Lap): three or four (182-221 VGPRs in FP64) leave 2 waves per SIMD;
    // on 96^3 x 64 functions H psi takes 0.28 / 0.37 / 0.46 ms with 5 / 10 /
    // 16 rows and the FP64 V-cycle 4.0 / 3.8 / 4.2 ms
    const int maxitems = mode == MG_HLOC ? 256 : 512;
    g.ty = std::min(std::min(16, std::max(1, maxitems / G)), d1);
    const size_t lds = 2 * (size_t)(g.ty + 2 * R) * (d2 + 4) * sizeof(T);
    if (lds > 64 * 1024) return 0;
    g.nty = (d1 + g.ty - 1) / g.ty;
    // enough blocks to fill the chip: split x into segments (each re-reads
    // 2R planes) when the functions and y tiles alone are too few
    const int64_t base = (int64_t)nfunc * g.nty;
    g.nxs  = (int)std::max<int64_t>(1, std::min<int64_t>((512 + base - 1) / base, std::max(1, d0 / 8)));
    g.xlen = (d0 + g.nxs - 1) / g.nxs;
    g.nxs  = (d0 + g.xlen - 1) / g.xlen;
    g.nreal = base * g.nxs;
    const int items = g.ty * G;
    const dim3 grid((unsigned)((g.nreal + 7) / 8 * 8));
    const Coef4 c4 = coef4(gr.h);
    const Coef2 c2 = coef2(gr.h);
    if (mode == MG_LAP)
    {
        if (type == MGX_LAP_4TH)
            march_grid_nt<T, MGX_LAP_4TH, MG_LAP, false>(ctx, items, grid, lds, in, f, vtot, out, w, g, c4, c2, mjf);
        else
            march_grid_nt<T, MGX_LAP_2ND, MG_LAP, false>(ctx, items, grid, lds, in, f, vtot, out, w, g, c4, c2, mjf);
    }
    else if (mode == MG_JACOBI)
    {
        if (type == MGX_LAP_4TH)
            march_grid_nt<T, MGX_LAP_4TH, MG_JACOBI, true>(ctx, items, grid, lds, in, f, vtot, out, w, g, c4, c2, mjf);
        else
            march_grid_nt<T, MGX_LAP_2ND, MG_JACOBI, true>(ctx, items, grid, lds, in, f, vtot, out, w, g, c4, c2, mjf);
    }
    else
        march_grid_nt<T, MGX_LAP_4TH, MG_HLOC, true>(ctx, items, grid, lds, in, f, vtot, out, w, g, c4, c2, mjf);
    MGX_CHECK_LAUNCH(ctx);
    return 1;
}

template <typename T>
int launch_lap(mgx_ctx* ctx, int type, const T* in, T* out, int nfunc, const mgx_grid_t& gr)
{
    if ((type == MGX_LAP_4TH && gr.ghosts >= 2) || (type == MGX_LAP_2ND && gr.ghosts >= 1))
    {
        const MView v = mview_ghosted(gr);
        const int m   = march_grid(ctx, type, MG_LAP, in, v, (const T*)nullptr, nullptr, out, v, (T*)nullptr, nfunc,
            gr, 0.);
        if (m < 0) return m;
        if (m) return MGX_OK;
    }
    GGeom g = ggeom(gr);
    AllCoef c;
    c.c4 = coef4(gr.h);
    c.c2 = coef2(gr.h);
    c.cm = coefm(gr.h);
    const int64_t total = (int64_t)nfunc * g.size;
    dim3 grd(nblocks(total)), blk(kBlock);
    switch (type)
    {
        case MGX_LAP_4TH:
            MGX_REQUIRE(ctx, gr.ghosts >= 2, "4th-order stencil needs 2 ghosts");
            hipLaunchKernelGGL((k_lap<T, MGX_LAP_4TH>), grd, blk, 0, ctx->stream, in, out, g, total, c);
            break;
        case MGX_LAP_2ND:
            MGX_REQUIRE(ctx, gr.ghosts >= 1, "2nd-order stencil needs 1 ghost");
            hipLaunchKernelGGL((k_lap<T, MGX_LAP_2ND>), grd, blk, 0, ctx->stream, in, out, g, total, c);
            break;
        case MGX_LAP_MEHRSTELLEN:
            MGX_REQUIRE(ctx, gr.ghosts >= 1, "Mehrstellen stencil needs 1 ghost");
            hipLaunchKernelGGL((k_lap<T, MGX_LAP_MEHRSTELLEN>), grd, blk, 0, ctx->stream, in, out, g, total, c);
            break;
        default:
            return set_err(ctx, MGX_EINVAL, "LapFactory::createLap() --- option invalid");
    }
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
int launch_axpy(mgx_ctx* ctx, int64_t n, double a, const T* x, T* y)
{
    hipLaunchKernelGGL(k_axpy<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, n, a, x, y);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// one weighted-Jacobi sweep with the result in xo when the marching kernel
// applies (*moved = true: no halo exchange of x is needed, and the ghost
// cells of xo and w are not written), else in place on x by launch_jacobi
template <typename T>
int launch_jacobi(mgx_ctx* ctx, int type, double jf, const T* f, T* x, T* w, int nfunc, const mgx_grid_t& gr);

template <typename T>
int jacobi_sweep(mgx_ctx* ctx, int type, double jf, const T* f, T* x, T* xo, T* w, int nfunc, const mgx_grid_t& gr,
    bool* moved, bool need_w = true)
{
    *moved = false;
    if (xo)
    {
        // the marching sweep writes w = A x - f only where the caller reads it
        const MView v = mview_ghosted(gr);
        const int m   = march_grid(ctx, type, MG_JACOBI, (const T*)x, v, f, nullptr, xo, v, need_w ? w : (T*)nullptr,
            nfunc, gr, -1. * jf);
        if (m < 0) return m;
        if (m)
        {
            *moved = true;
            return MGX_OK;
        }
        if (type == MGX_LAP_MEHRSTELLEN && gr.ghosts >= 1)
        {
            // the Mehrstellen stencil and the update in one pass (after the
            // halo exchange its ghosts need)
            int rc = launch_halo(ctx, x, nfunc, gr);
            if (rc) return rc;
            const GGeom g = ggeom(gr);
            AllCoef c;
            c.c4 = coef4(gr.h);
            c.c2 = coef2(gr.h);
            c.cm = coefm(gr.h);
            const int64_t total = (int64_t)nfunc * g.size;
            hipLaunchKernelGGL(k_jacobi_mehr<T>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, (const T*)x, f,
                xo, need_w ? w : (T*)nullptr, g, total, c, -1. * jf);
            MGX_CHECK_LAUNCH(ctx);
            *moved = true;
            return MGX_OK;
        }
    }
    return launch_jacobi(ctx, type, jf, f, x, w, nfunc, gr);
}

template <typename T>
int launch_jacobi(mgx_ctx* ctx, int type, double jf, const T* f, T* x, T* w, int nfunc, const mgx_grid_t& gr)
{
    int rc = launch_halo(ctx, x, nfunc, gr);
    if (rc) return rc;
    rc = launch_lap(ctx, type, x, w, nfunc, gr);
    if (rc) return rc;
    const int64_t n = (int64_t)nfunc * ggeom(gr).sizeg;
    hipLaunchKernelGGL(k_jacobi_update<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, n, -1. * jf, f, w, x);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// GridFuncVector::jacobi on one block (mgx_jacobi): the marching sweep into a
// workspace and the interior copied back into x (the ghost cells of x and w
// keep what they held: every consumer exchanges halos first), else the
// halo + stencil + update kernels in place
template <typename T>
int jacobi_entry(mgx_ctx* ctx, int type, double jf, const T* f, T* x, T* w, int nfunc, const mgx_grid_t& gr)
{
    const GGeom g = ggeom(gr);
    T* xo         = (T*)workspace(ctx, 1, (size_t)nfunc * g.sizeg * sizeof(T));
    if (!xo) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    bool moved = false;
    int rc     = jacobi_sweep(ctx, type, jf, f, x, xo, w, nfunc, gr, &moved);
    if (rc || !moved) return rc;
    const int64_t total = (int64_t)nfunc * g.size;
    hipLaunchKernelGGL(k_copy_interior<T>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, (const T*)xo, x, g,
        total);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
int launch_restrict(mgx_ctx* ctx, const T* uf, const mgx_grid_t& gf, T* uc, const mgx_grid_t& gc, int nfunc)
{
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, gf.dim[i] >= 2 * gc.dim[i], "coarse grid too large");
    MGX_REQUIRE(ctx, gf.ghosts >= 1, "restriction needs fine ghosts");
    GGeom a = ggeom(gf), b = ggeom(gc);
    const int64_t total = (int64_t)nfunc * b.size;
    hipLaunchKernelGGL(k_restrict<T>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, uf, a, uc, b, total);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
int launch_extend(mgx_ctx* ctx, const T* uc, const mgx_grid_t& gc, T* uf, const mgx_grid_t& gf, int nfunc)
{
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, gf.dim[i] >= 2 * gc.dim[i], "coarse grid too large");
    MGX_REQUIRE(ctx, gf.ghosts >= 1 && gc.ghosts >= 1, "prolongation needs ghosts");
    GGeom a = ggeom(gc), b = ggeom(gf);
    const int64_t ninj = (int64_t)nfunc * (gc.dim[0] + 1) * (gc.dim[1] + 1) * (gc.dim[2] + 1);
    hipLaunchKernelGGL(k_inject<T>, dim3(nblocks(ninj)), dim3(kBlock), 0, ctx->stream, uc, a, uf, b, ninj);
    MGX_CHECK_LAUNCH(ctx);
    const int64_t total = (int64_t)nfunc * b.size;
    hipLaunchKernelGGL(k_fill_odd<T>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, uf, b, total);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// x -= mask * P(uc) on the fine interior (uc with fresh ghosts), in place of
// launch_extend + launch_mask + launch_axpy
template <typename T>
int launch_prolong_sub(mgx_ctx* ctx, const T* uc, const mgx_grid_t& gc, T* xf, const mgx_grid_t& gf, int nfunc,
    const uint8_t* mask)
{
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, gf.dim[i] == 2 * gc.dim[i], "fine grid is not twice the coarse grid");
    MGX_REQUIRE(ctx, gf.ghosts >= 1 && gc.ghosts >= 1, "prolongation needs ghosts");
    const GGeom a = ggeom(gc), b = ggeom(gf);
    const int np = gf.dim[1] * gf.dim[2];
    const dim3 grid((unsigned)((np + 255) / 256), (unsigned)(nfunc * gf.dim[0]));
    hipLaunchKernelGGL(k_prolong_sub<T>, grid, dim3(256), 0, ctx->stream, uc, a, xf, b, mask);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
int launch_mask(mgx_ctx* ctx, T* u, const uint8_t* m, int nfunc, const mgx_grid_t& gr)
{
    if (!m) return MGX_OK;
    GGeom g             = ggeom(gr);
    const int64_t total = (int64_t)nfunc * g.size;
    hipLaunchKernelGGL(k_mask<T>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, u, m, g, total);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
int launch_fill(mgx_ctx* ctx, T* u, int64_t n, T v)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, u, n, v);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

mgx_grid_t coarse_of(const mgx_grid_t& g)
{
    // Grid::coarse_grid, src/pb/Grid.cc:214-231
    mgx_grid_t c = g;
    for (int i = 0; i < 3; i++)
    {
        c.dim[i] = g.dim[i] / 2;
        c.h[i]   = 2. * g.h[i];
    }
    return c;
}

template <typename T>
struct MGBufs
{
    std::vector<mgx_grid_t> grids;
    std::vector<T*> w, t, rc, nv;  // t: the other buffer of the out-of-place sweeps
};

template <typename T>
int mg_setup(mgx_ctx* ctx, MGBufs<T>& B, const mgx_grid_t& g0, int max_levels, int nfunc, int slot)
{
    MGX_REQUIRE(ctx, max_levels >= 0 && max_levels < 10, "bad number of levels");
    B.grids.push_back(g0);
    for (int l = 1; l <= max_levels; l++)
    {
        const mgx_grid_t& p = B.grids.back();
        for (int i = 0; i < 3; i++)
            MGX_REQUIRE(ctx, p.dim[i] % 2 == 0, "grid not divisible by 2 for coarsening");
        B.grids.push_back(coarse_of(p));
    }
    size_t total = 0;
    std::vector<size_t> sz(max_levels + 1);
    for (int l = 0; l <= max_levels; l++)
    {
        sz[l] = (size_t)nfunc * ggeom(B.grids[l]).sizeg;
        total += 2 * sz[l];
        if (l > 0) total += 2 * sz[l];
    }
    T* base = (T*)workspace(ctx, slot, total * sizeof(T));
    if (!base) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    T* p = base;
    for (int l = 0; l <= max_levels; l++)
    {
        B.w.push_back(p);
        p += sz[l];
        B.t.push_back(p);
        p += sz[l];
    }
    for (int l = 1; l <= max_levels; l++)
    {
        B.rc.push_back(p);
        p += sz[l];
        B.nv.push_back(p);
        p += sz[l];
    }
    return MGX_OK;
}

// Preconditioning<T>::mg, src/Preconditioning.cc:155-216
template <typename T>
// pout (FP64 V-cycles, no masks, periodic grids): the last sweep of the call
// writes the packed result there (*pout_done = true; v then keeps the
// previous iterate) when the marching kernel applies
int mg_rec(mgx_ctx* ctx, MGBufs<T>& B, int max_levels, T* v, const T* f, int type, int level, int nfunc,
    const uint8_t* const* masks, double* pout = nullptr, int64_t ldo = 0, bool* pout_done = nullptr)
{
    const mgx_grid_t& g = B.grids[level];
    const int ncycl     = (level == max_levels) ? 4 : 2;
    const double jf     = jacobi_factor(type, g.h);
    T* w                = B.w[level];
    const uint8_t* mk   = masks ? masks[level] : nullptr;
    const size_t bytes  = (size_t)nfunc * ggeom(g).sizeg * sizeof(T);
    int rc;
    // the sweeps alternate between v and t when they run out of place
    T* cur     = v;
    // w is read after the last pre-smoothing sweep only (the stale residual
    // the coarse correction restricts)
    auto sweep = [&](bool need_w) -> int {
        bool moved = false;
        T* other   = cur == v ? B.t[level] : v;
        int r      = jacobi_sweep(ctx, type, jf, f, cur, other, w, nfunc, g, &moved, need_w);
        if (r) return r;
        if (moved) cur = other;
        return launch_mask(ctx, cur, mk, nfunc, g);
    };
    auto settle = [&]() -> int {
        if (cur != v) MGX_CHECK_HIP(ctx, hipMemcpyAsync(v, cur, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        cur = v;
        return MGX_OK;
    };
    for (int it = 0; it < ncycl; it++)
        if ((rc = sweep(it == ncycl - 1 && level < max_levels))) return rc;
    if ((rc = settle())) return rc;
    if (level == max_levels) return MGX_OK;
    if ((rc = launch_mask(ctx, w, mk, nfunc, g))) return rc;
    const mgx_grid_t& gc = B.grids[level + 1];
    T* rcoarse           = B.rc[level];
    if ((rc = launch_halo(ctx, w, nfunc, g))) return rc;
    if ((rc = launch_restrict(ctx, w, g, rcoarse, gc, nfunc))) return rc;
    if ((rc = launch_mask(ctx, rcoarse, masks ? masks[level + 1] : nullptr, nfunc, gc))) return rc;
    T* nv = B.nv[level];
    if ((rc = launch_fill(ctx, nv, (int64_t)nfunc * ggeom(gc).sizeg, (T)0))) return rc;
    if ((rc = mg_rec(ctx, B, max_levels, nv, rcoarse, MGX_LAP_2ND, level + 1, nfunc, masks))) return rc;
    if ((rc = launch_halo(ctx, nv, nfunc, gc))) return rc;
    if ((rc = launch_prolong_sub(ctx, (const T*)nv, gc, v, g, nfunc, mk))) return rc;
    for (int it = 0; it < 2; it++)
    {
        if constexpr (std::is_same<T, double>::value)
            if (it == 1 && pout && !mk && g.bc[0] == 1 && g.bc[1] == 1 && g.bc[2] == 1)
            {
                const int m = march_grid(ctx, type, MG_JACOBI, (const T*)cur, mview_ghosted(g), f, nullptr, pout,
                    mview_packed(g, ldo), (T*)nullptr, nfunc, g, -1. * jf);
                if (m < 0) return m;
                if (m)
                {
                    *pout_done = true;
                    return MGX_OK;
                }
            }
        if ((rc = sweep(false))) return rc;
    }
    if ((rc = settle())) return rc;
    if (g.bc[0] != 1 || g.bc[1] != 1 || g.bc[2] != 1) return launch_halo(ctx, v, nfunc, g);
    return MGX_OK;
}

template <typename T>
int mg_run(mgx_ctx* ctx, int type, int max_levels, T* v, const T* f, int nfunc, const mgx_grid_t& g0,
    const uint8_t* const* masks, int slot, double* pout = nullptr, int64_t ldo = 0, bool* pout_done = nullptr)
{
    MGX_REQUIRE(ctx, type >= 0 && type <= 2, "LapFactory::createLap() --- option invalid");
    MGBufs<T> B;
    int rc = mg_setup(ctx, B, g0, max_levels, nfunc, slot);
    if (rc) return rc;
    return mg_rec(ctx, B, max_levels, v, f, type, 0, nfunc, masks, pout, ldo, pout_done);
}

template <typename T>
int precond_run(mgx_ctx* ctx, int type, int max_levels, double gamma, const mgx_grid_t& g0, const double* res,
    int64_t ld, int nfunc, const uint8_t* const* masks, double* out, int64_t ldo)
{
    GGeom g        = ggeom(g0);
    const int64_t n = (int64_t)nfunc * g.sizeg;
    T* buf         = (T*)workspace(ctx, 0, 2 * n * sizeof(T));
    if (!buf) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    T* x = buf;
    T* f = buf + n;
    // f from the packed residual and the start x = gamma f in one pass
    hipLaunchKernelGGL(k_packed_to_ghosted_start<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, res, ld, f, x,
        gamma, g, n);
    MGX_CHECK_LAUNCH(ctx);
    // the last fine sweep may write the packed output itself
    bool done = false;
    int rc    = mg_run(ctx, type, max_levels, x, (const T*)f, nfunc, g0, masks, 1, out, ldo, &done);
    if (rc || done) return rc;
    const int64_t m = (int64_t)nfunc * g.size;
    hipLaunchKernelGGL(k_ghosted_to_packed<T>, dim3(nblocks(m)), dim3(kBlock), 0, ctx->stream, x, g, out, ldo, m);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// ------------------------------------------------- Hartree multigrid ----
// vel * sum of squares (or the plain sum) over the interior of one function:
// per-block partials, then one block adds them in a fixed order
template <bool SQ>
__global__ void __launch_bounds__(256) k_int_sum(const double* u, GGeom g, int lo0, int lo1, int lo2, double* part)
{
    __shared__ double lds[4];
    double acc = 0.;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        int r       = (int)t;
        const int x = r / (g.d1 * g.d2);
        r -= x * g.d1 * g.d2;
        const int y  = r / g.d2;
        const int z  = r - y * g.d2;
        if (x < lo0 || y < lo1 || z < lo2) continue;
        const double v = u[(int64_t)(x + g.ng) * g.incx + (int64_t)(y + g.ng) * g.incy + z + g.ng];
        acc += SQ ? v * v : v;
    }
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

__global__ void __launch_bounds__(256) k_part_sum(const double* part, int n, double* out)
{
    __shared__ double lds[4];
    double acc = 0.;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        acc += part[i];
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

// c = a - b (GridFunc::diff, src/pb/GridFunc.cc:557-572), all sizeg points
__global__ void k_diff(const double* a, const double* b, double* c, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        c[t] = a[t] - b[t];
}

__global__ void k_add_const(double* u, int64_t n, const double* s, double scale)
{
    const double v = -scale * s[0];
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        u[t] += v;
}

// ------------------------------------------- Poisson-Boltzmann (PBh4M) --
// interior point t of a grid -> offset in the ghosted layout
__device__ __forceinline__ int64_t gpoint(const GGeom& g, int64_t t, int* xo = nullptr, int* yo = nullptr,
    int* zo = nullptr)
{
    int r       = (int)t;
    const int x = r / (g.d1 * g.d2);
    r -= x * g.d1 * g.d2;
    const int y = r / g.d2;
    const int z = r - y * g.d2;
    if (xo)
    {
        *xo = x;
        *yo = y;
        *zo = z;
    }
    return (int64_t)(x + g.ng) * g.incx + (int64_t)(y + g.ng) * g.incy + z + g.ng;
}

// DielFunc::epsilon_rho (src/pb/DielFunc.cc:38-43) at every interior point,
// then sqrt_a = sqrt(eps) and inv_sqrt_a = 1 / sqrt_a (GridFunc::sqrt_func,
// GridFunc::inv)
__device__ __forceinline__ double pb_eps_of(double r, double e0, double rho0, double drho0)
{
    const double ratio = pow(r / rho0, 2. * drho0);
    return 1. + 0.5 * (e0 - 1.) * ((1. - ratio) / (1. + ratio) + 1.);
}

__global__ void k_pb_eps(const double* rhod, GGeom g, double e0, double rho0, double drho0, double* sa, double* isa)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        const double s  = sqrt(pb_eps_of(rhod[i], e0, rho0, drho0));
        sa[i]           = s;
        isa[i]          = 1. / s;
    }
}

// PBh4<T>::init / PBh2 (DielFunc::Gepsilon_rho): eps at every interior point
__global__ void k_pbg_eps(const double* rhod, GGeom g, double e0, double rho0, double drho0, double* eps)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        eps[i]          = pb_eps_of(rhod[i], e0, rho0, drho0);
    }
}

// the first difference of FDoper::del1_4th (ORDER 4, src/pb/FDoper.cc:119-178)
// or del1_2nd (ORDER 2, :72-115) along the axis of stride c, ih = 1 / h
template <int ORDER>
__device__ __forceinline__ double pb_d1(const double* u, int64_t c, double ih)
{
    if constexpr (ORDER == 4)
    {
        const double e1 = (8. / 12.) * ih, e2 = (1. / 12.) * ih;
        return e1 * (u[c] - u[-c]) + e2 * (u[-2 * c] - u[2 * c]);
    }
    else
        return (0.5 * ih) * (u[c] - u[-c]);
}

// the gradient form of -div(eps grad x), PBh4<T>::pb_4th (src/pb/PBh4.cc:13-28)
// and PBh2<T>::pb_2nd (src/pb/PBh2.cc:13-30): out holds -Lap x on entry;
// out = (((0 - dx0 de0) - dx1 de1) - dx2 de2) + (-Lap x) eps, x and eps with
// fresh ghosts
template <int ORDER>
__global__ void k_pbg_grad(const double* x, const double* eps, GGeom g, double ih0, double ih1, double ih2,
    double* out)
{
    const int64_t inc[3] = { g.incx, g.incy, 1 };
    const double ih[3]   = { ih0, ih1, ih2 };
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        double acc      = 0.;
#pragma unroll
        for (int d = 0; d < 3; d++)
            acc = acc - pb_d1<ORDER>(x + i, inc[d], ih[d]) * pb_d1<ORDER>(eps + i, inc[d], ih[d]);
        out[i] = acc + out[i] * eps[i];
    }
}

// PBh4<T>::jacobi / PBh2<T>::jacobi (GridFunc<T>::jacobi,
// src/pb/GridFunc.cc:3264-3280): x -= (1 / (c0 eps)) (w - f), w = A x
__global__ void k_pbg_jacobi(double* x, const double* w, const double* f, const double* eps, GGeom g, double c0)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        x[i] -= (1. / (c0 * eps[i])) * (w[i] - f[i]);
    }
}

// pp = (-1 * L) * inv_sqrt_a, L = -Lap_4 sqrt_a (PBh4M<T>::init)
__global__ void k_pb_pp(const double* L, const double* isa, GGeom g, double* pp)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        pp[i]           = (-1. * L[i]) * isa[i];
    }
}

// out += B(pp x) (PBh4M<T>::apply: the product on the 7 points, then
// FDkernelRHS_4th_Mehr1's arithmetic as in k_rhs); x and pp with fresh ghosts
__global__ void k_pb_bterm(const double* x, const double* pp, GGeom g, double* out)
{
    const int64_t X = g.incx, Y = g.incy;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        const double* a = x + i;
        const double* q = pp + i;
        const double b  = 0.5 * (q[0] * a[0])
                         + kInv12 * ((((((q[-X] * a[-X]) + (q[X] * a[X])) + (q[-Y] * a[-Y])) + (q[Y] * a[Y]))
                                        + (q[-1] * a[-1]))
                                       + (q[1] * a[1]));
        out[i] = out[i] + b;
    }
}

// PBh4M<T>::jacobi update: x -= (0.75 / (c0 + pp)) (w - f), w = A x
__global__ void k_pb_jacobi(double* x, const double* w, const double* f, const double* pp, GGeom g, double c0)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        x[i] -= (0.75 / (c0 + pp[i])) * (w[i] - f[i]);
    }
}

// interior out = a * b (mode 0), a / b (mode 1) or a (mode 2)
__global__ void k_pb_scale(const double* a, const double* b, GGeom g, int mode, double* out)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        out[i]          = mode == 0 ? a[i] * b[i] : (mode == 1 ? a[i] / b[i] : a[i]);
    }
}

// get_vepsilon of PBh4M / PBh4 (src/pb/PBh4M.cc:138-162, src/pb/PBh4.cc:49-63)
// and PBh2 (src/pb/PBh2.cc:57-72): vks = depsilon(rhod) |grad vh|^2 (-1/(8 pi)),
// the gradient by del1_4th (ORDER 4) or del1_2nd (ORDER 2); vh with fresh ghosts
template <int ORDER>
__global__ void k_pb_vks(const double* vh, const double* rhod, GGeom g, double h0, double h1, double h2, double e0,
    double rho0, double drho0, double* out)
{
    const int64_t inc[3] = { g.incx, g.incy, 1 };
    const double hh[3]   = { h0, h1, h2 };
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.size; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t i = gpoint(g, t);
        double g2       = 0.;
        for (int d = 0; d < 3; d++)
        {
            const double dv = pb_d1<ORDER>(vh + i, inc[d], 1. / hh[d]);
            g2 += dv * dv;
        }
        const double r = rhod[i];
        double de      = 0.;
        if (!(r < 0.001 * rho0))
        {
            const double inv = 1. / rho0, ratio = r * inv, pr = pow(ratio, 2. * drho0);
            de = (2. * inv) * (1. - e0) * drho0 * pow(ratio, 2. * drho0 - 1.) / ((1. + pr) * (1. + pr));
        }
        out[i] = (de * g2) * (-1. / (8. * M_PI));
    }
}

struct PoissonMG
{
    std::vector<mgx_grid_t> grids;
    std::vector<int> types;
    std::vector<double*> res, x, f, t;  // t: the other buffer of the out-of-place sweeps
    double* part = nullptr;  // kPartBlocks partials + 8 results
};

constexpr int kPartBlocks = 1024;

// one sum over the interior to dev_out (device); SQ: of squares.  As in
// GridFunc::gdot (src/pb/GridFunc.cc:2745-2800), the index-0 layer of a
// Dirichlet direction (the boundary plane) is left out.
template <bool SQ>
int int_sum(mgx_ctx* ctx, PoissonMG& P, const double* u, const mgx_grid_t& gr, double* dev_out)
{
    GGeom g = ggeom(gr);
    const int nb = (int)std::min<int64_t>(kPartBlocks, (g.size + 255) / 256);
    hipLaunchKernelGGL(k_int_sum<SQ>, dim3(nb), dim3(256), 0, ctx->stream, u, g, gr.bc[0] != 1, gr.bc[1] != 1,
        gr.bc[2] != 1, P.part);
    hipLaunchKernelGGL(k_part_sum, dim3(1), dim3(256), 0, ctx->stream, (const double*)P.part, nb, dev_out);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// pb::norm: sqrt(vel * sum u^2) over the interior (host result)
int int_norm(mgx_ctx* ctx, PoissonMG& P, const double* u, const mgx_grid_t& gr, double* out)
{
    int rc = int_sum<true>(ctx, P, u, gr, P.part + kPartBlocks);
    if (rc) return rc;
    double v = 0.;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(&v, P.part + kPartBlocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = std::sqrt(v * gr.h[0] * gr.h[1] * gr.h[2]);
    return MGX_OK;
}

// pb::Vcycle, src/pb/Vcycle.h:28-249 (one task, USE_LOWER_ORDER); every level
// keeps the fine grid's boundary conditions
int poisson_vcycle(mgx_ctx* ctx, PoissonMG& P, int l, double* x, const double* f, int nu1, int nu2)
{
    const mgx_grid_t& g = P.grids[l];
    const int t         = P.types[l];
    const double jf     = jacobi_factor(t, g.h);
    double* res         = P.res[l];
    const int64_t n     = ggeom(g).sizeg;
    int rc;
    // the sweeps alternate between x and P.t[l] when they run out of place
    double* cur = x;
    auto sweeps = [&](int k) -> int {
        for (int i = 0; i < k; i++)
        {
            bool moved    = false;
            double* other = cur == x ? P.t[l] : x;
            // the fresh residual is formed after the sweeps: w is not read
            int r         = jacobi_sweep(ctx, t, jf, f, cur, other, res, 1, g, &moved, false);
            if (r) return r;
            if (moved) cur = other;
        }
        if (cur != x)
            MGX_CHECK_HIP(ctx, hipMemcpyAsync(x, cur, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        cur = x;
        return MGX_OK;
    };
    if ((rc = sweeps(nu1))) return rc;
    if (l + 1 < (int)P.grids.size())
    {
        const mgx_grid_t& gc = P.grids[l + 1];
        // fresh residual A x - rhs, restricted
        if ((rc = launch_halo(ctx, x, 1, g))) return rc;
        if ((rc = launch_lap(ctx, t, (const double*)x, res, 1, g))) return rc;
        if ((rc = launch_axpy(ctx, n, -1., f, res))) return rc;
        if ((rc = launch_halo(ctx, res, 1, g))) return rc;
        if ((rc = launch_restrict(ctx, (const double*)res, g, P.f[l + 1], gc, 1))) return rc;
        // coarse correction from zero, prolonged and subtracted
        if ((rc = launch_fill(ctx, P.x[l + 1], ggeom(gc).sizeg, 0.))) return rc;
        if ((rc = poisson_vcycle(ctx, P, l + 1, P.x[l + 1], P.f[l + 1], nu1, nu2))) return rc;
        if ((rc = launch_halo(ctx, P.x[l + 1], 1, gc))) return rc;
        if ((rc = launch_prolong_sub(ctx, (const double*)P.x[l + 1], gc, x, g, 1, (const uint8_t*)nullptr)))
            return rc;
    }
    return sweeps(nu2);
}

// pb::Mgm, src/pb/Mgm.h:12-111, as SolverLap::solve calls it
// (src/pb/SolverLap.cc:49-63); Dirichlet directions hold zero boundary values
int poisson_run(mgx_ctx* ctx, int type, const mgx_grid_t& g0, const double* rho, double* vh, int nu1, int nu2,
    int max_sweeps, double tol, int max_nlevels, mgx_poisson_stats_t* st)
{
    PoissonMG P;
    P.grids.push_back(g0);
    P.types.push_back(type);
    // Vcycle coarsens level l while l < max_nlevels and every dimension is
    // even and at least twice the ghost width; coarse grids keep one ghost
    // and the lower-order operator (Laph4 -> Laph2)
    while ((int)P.grids.size() - 1 < max_nlevels)
    {
        const mgx_grid_t& p = P.grids.back();
        bool ok = true;
        for (int i = 0; i < 3; i++)
            ok = ok && p.dim[i] % 2 == 0 && p.dim[i] >= 2 * p.ghosts;
        if (!ok) break;
        mgx_grid_t c = coarse_of(p);
        c.ghosts     = 1;
        P.grids.push_back(c);
        P.types.push_back(type == MGX_LAP_4TH ? MGX_LAP_2ND : type);
    }
    const int nl = (int)P.grids.size();
    const int64_t n0 = ggeom(g0).sizeg;
    size_t total = 4 * (size_t)n0 + kPartBlocks + 8;  // rhs, lhs, r0 (then res0), w
    for (int l = 0; l < nl; l++)
        total += (size_t)ggeom(P.grids[l]).sizeg * (l == 0 ? 2 : 4);
    double* base = (double*)workspace(ctx, 0, total * sizeof(double));
    if (!base) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    double* rhs = base;
    double* lhs = rhs + n0;
    double* r0  = lhs + n0;
    double* w   = r0 + n0;
    double* q   = w + n0;
    P.part      = q;
    q += kPartBlocks + 8;
    for (int l = 0; l < nl; l++)
    {
        const int64_t m = ggeom(P.grids[l]).sizeg;
        P.res.push_back(q);
        q += m;
        P.t.push_back(q);
        q += m;
        if (l == 0)
        {
            P.x.push_back(nullptr);
            P.f.push_back(nullptr);
            continue;
        }
        P.x.push_back(q);
        q += m;
        P.f.push_back(q);
        q += m;
    }
    int rc;
    // rhs = A.rhs(rho): B rho (Mehrstellen), rho otherwise
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(r0, rho, n0 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if ((rc = launch_halo(ctx, r0, 1, g0))) return rc;
    if (type == MGX_LAP_MEHRSTELLEN)
    {
        GGeom a = ggeom(g0);
        hipLaunchKernelGGL(k_rhs<double>, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, (const double*)r0, rhs,
            a, a, a.size);
        MGX_CHECK_LAUNCH(ctx);
    }
    else
        MGX_CHECK_HIP(ctx, hipMemcpyAsync(rhs, r0, n0 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    double rhs_norm = 0.;
    if ((rc = int_norm(ctx, P, rhs, g0, &rhs_norm))) return rc;
    MGX_REQUIRE(ctx, rhs_norm > 0., "zero right-hand side");
    const double inv_rhs_norm = 1. / rhs_norm;
    double init = 1., fin = 1., rel = 1.;
    int conv = 0, sweeps = 0;
    for (int i = 0; i < max_sweeps; i++)
    {
        if ((rc = launch_halo(ctx, vh, 1, g0))) return rc;
        if ((rc = launch_lap(ctx, type, (const double*)vh, lhs, 1, g0))) return rc;
        hipLaunchKernelGGL(k_diff, dim3(nblocks(n0)), dim3(kBlock), 0, ctx->stream, (const double*)rhs,
            (const double*)lhs, r0, n0);
        MGX_CHECK_LAUNCH(ctx);
        double rn = 0.;
        if ((rc = int_norm(ctx, P, r0, g0, &rn))) return rc;
        if (i == 0) init = rn;
        if (rn * inv_rhs_norm < tol)
        {
            fin  = rn;
            rel  = rn * inv_rhs_norm;
            conv = 1;
            break;
        }
        if ((rc = launch_fill(ctx, w, n0, 0.))) return rc;
        if ((rc = poisson_vcycle(ctx, P, 0, w, r0, nu1, nu2))) return rc;
        sweeps++;
        if ((rc = launch_axpy(ctx, n0, 1., (const double*)w, vh))) return rc;
    }
    if (!conv)
    {
        if ((rc = launch_halo(ctx, vh, 1, g0))) return rc;
        if ((rc = launch_lap(ctx, type, (const double*)vh, lhs, 1, g0))) return rc;
        if ((rc = launch_axpy(ctx, n0, -1., (const double*)rhs, lhs))) return rc;
        if ((rc = int_norm(ctx, P, lhs, g0, &fin))) return rc;
        rel = fin * inv_rhs_norm;
    }
    // GridFunc::average0 on a fully periodic grid only (SolverLap::solve,
    // src/pb/SolverLap.cc:60-62), then fresh ghosts
    if (g0.bc[0] == 1 && g0.bc[1] == 1 && g0.bc[2] == 1)
    {
        GGeom a = ggeom(g0);
        if ((rc = int_sum<false>(ctx, P, vh, g0, P.part + kPartBlocks))) return rc;
        hipLaunchKernelGGL(k_add_const, dim3(nblocks(n0)), dim3(kBlock), 0, ctx->stream, vh, n0,
            (const double*)(P.part + kPartBlocks), 1. / (double)a.size);
        MGX_CHECK_LAUNCH(ctx);
    }
    if ((rc = launch_halo(ctx, vh, 1, g0))) return rc;
    if (st)
    {
        st->converged               = conv;
        st->nb_sweeps               = sweeps;
        st->final_residual          = fin;
        st->final_relative_residual = rel;
        st->residual_reduction      = fin / init;
    }
    return MGX_OK;
}

// SolverPB<PBh4M / PBh4 / PBh2>::solve (src/pb/SolverPB.cc:56-84): the
// dielectric operator's Mgm (on the transformed potential for PBh4M), then
// get_vepsilon.  op per level: MGX_LAP_MEHRSTELLEN (PBh4M, coef = pp),
// MGX_LAP_4TH (PBh4, coef = eps) or MGX_LAP_2ND (PBh2, coef = eps)
struct PBLevel
{
    mgx_grid_t g;
    int op;
    double *pp, *x, *f, *w, *t;
};

// the dielectric operator at level l (x gets its halo): PBh4M, out =
// -Lap_M x + B(pp x); PBh4 / PBh2, out = eps (-Lap x) - grad x . grad eps
int pb_apply(mgx_ctx* ctx, const PBLevel& L, double* x, double* out)
{
    int rc;
    if ((rc = launch_halo(ctx, x, 1, L.g))) return rc;
    if ((rc = launch_lap(ctx, L.op, (const double*)x, out, 1, L.g))) return rc;
    const GGeom a = ggeom(L.g);
    if (L.op == MGX_LAP_MEHRSTELLEN)
        hipLaunchKernelGGL(k_pb_bterm, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, (const double*)x,
            (const double*)L.pp, a, out);
    else
        hipLaunchKernelGGL(L.op == MGX_LAP_4TH ? k_pbg_grad<4> : k_pbg_grad<2>, dim3(nblocks(a.size)), dim3(kBlock), 0,
            ctx->stream, (const double*)x, (const double*)L.pp, a, 1. / L.g.h[0], 1. / L.g.h[1], 1. / L.g.h[2], out);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int pb_vcycle(mgx_ctx* ctx, std::vector<PBLevel>& Lv, PoissonMG& P, int l, double* x, const double* f, int nu1,
    int nu2)
{
    const PBLevel& L  = Lv[l];
    const GGeom a     = ggeom(L.g);
    const double ih2  = 1. / (L.g.h[0] * L.g.h[0]) + 1. / (L.g.h[1] * L.g.h[1]) + 1. / (L.g.h[2] * L.g.h[2]);
    const bool mehr   = L.op == MGX_LAP_MEHRSTELLEN;
    // PBh4M: (30/12) sum 1/h^2 (+ pp in the kernel); PBh4, PBh2: 1.5 (30/12) sum 1/h^2
    const double c0   = mehr ? (30. / 12.) * ih2 : 1.5 * (30. / 12.) * ih2;
    int rc;
    auto sweeps = [&](int k) -> int {
        for (int i = 0; i < k; i++)
        {
            int r = pb_apply(ctx, L, x, L.w);
            if (r) return r;
            hipLaunchKernelGGL(mehr ? k_pb_jacobi : k_pbg_jacobi, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream,
                x, (const double*)L.w, f, (const double*)L.pp, a, c0);
            MGX_CHECK_LAUNCH(ctx);
        }
        return MGX_OK;
    };
    if ((rc = sweeps(nu1))) return rc;
    if (l + 1 < (int)Lv.size())
    {
        const PBLevel& C = Lv[l + 1];
        if ((rc = pb_apply(ctx, L, x, L.w))) return rc;
        hipLaunchKernelGGL(k_diff, dim3(nblocks(a.sizeg)), dim3(kBlock), 0, ctx->stream, (const double*)L.w, f, L.t,
            a.sizeg);
        MGX_CHECK_LAUNCH(ctx);
        if ((rc = launch_halo(ctx, L.t, 1, L.g))) return rc;
        if ((rc = launch_restrict(ctx, (const double*)L.t, L.g, C.f, C.g, 1))) return rc;
        if ((rc = launch_fill(ctx, C.x, ggeom(C.g).sizeg, 0.))) return rc;
        if ((rc = pb_vcycle(ctx, Lv, P, l + 1, C.x, C.f, nu1, nu2))) return rc;
        if ((rc = launch_halo(ctx, C.x, 1, C.g))) return rc;
        if ((rc = launch_prolong_sub(ctx, (const double*)C.x, C.g, x, L.g, 1, (const uint8_t*)nullptr))) return rc;
    }
    return sweeps(nu2);
}

int pb_run(mgx_ctx* ctx, int op, const mgx_grid_t& g0, const double* rho, const double* rhod, double* vh,
    double* vks, double e0, double rho0, double drho0, int nu1, int nu2, int max_sweeps, double tol, int max_nlevels,
    mgx_poisson_stats_t* st)
{
    const bool mehr = op == MGX_LAP_MEHRSTELLEN;
    // levels: the fine grid, then while coarsenable (every dimension even and
    // at least twice the ghost width) grids of half the points with one ghost
    // (PBh4 and PBh2 take PBh2 below the fine level: USE_LOWER_ORDER,
    // PBh4<T>::getLowerOrderOp, src/pb/PBh4.h:72-80)
    std::vector<PBLevel> Lv(1);
    Lv[0].g  = g0;
    Lv[0].op = op;
    while ((int)Lv.size() - 1 < max_nlevels)
    {
        const mgx_grid_t& p = Lv.back().g;
        bool ok = true;
        for (int i = 0; i < 3; i++)
            ok = ok && p.dim[i] % 2 == 0 && p.dim[i] >= 2 * p.ghosts;
        if (!ok) break;
        PBLevel c;
        c.g        = coarse_of(p);
        c.g.ghosts = 1;
        c.op       = mehr ? MGX_LAP_MEHRSTELLEN : MGX_LAP_2ND;
        Lv.push_back(c);
    }
    const int nl     = (int)Lv.size();
    const int64_t n0 = ggeom(g0).sizeg;
    // fine: sa, isa, rhs, r0, corr + the level arrays (pp, w, t; x, f coarse)
    size_t total = 5 * (size_t)n0 + kPartBlocks + 8;
    for (int l = 0; l < nl; l++)
        total += (size_t)ggeom(Lv[l].g).sizeg * (l == 0 ? 3 : 5);
    double* base = (double*)workspace(ctx, 0, total * sizeof(double));
    if (!base) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    double* sa   = base;
    double* isa  = sa + n0;
    double* rhs  = isa + n0;
    double* r0   = rhs + n0;
    double* corr = r0 + n0;
    double* q    = corr + n0;
    PoissonMG P;
    P.part = q;
    q += kPartBlocks + 8;
    for (int l = 0; l < nl; l++)
    {
        const int64_t m = ggeom(Lv[l].g).sizeg;
        Lv[l].pp        = q;
        q += m;
        Lv[l].w = q;
        q += m;
        Lv[l].t = q;
        q += m;
        if (l == 0)
        {
            Lv[l].x = Lv[l].f = nullptr;
            continue;
        }
        Lv[l].x = q;
        q += m;
        Lv[l].f = q;
        q += m;
    }
    // the whole workspace starts at zero (ghosts the kernels do not write)
    if (int rc = launch_fill(ctx, base, (int64_t)total, 0.)) return rc;
    mgx_grid_t gp = g0;  // the dielectric functions are periodic
    gp.bc[0] = gp.bc[1] = gp.bc[2] = 1;
    const GGeom a = ggeom(g0);
    int rc;
    if (mehr)
    {
        // PBh4M<T>::init: eps, sqrt_a, inv_sqrt_a, pp = (-(-Lap_4 sqrt_a)) inv_sqrt_a
        hipLaunchKernelGGL(k_pb_eps, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, rhod, a, e0, rho0, drho0, sa,
            isa);
        MGX_CHECK_LAUNCH(ctx);
        if ((rc = launch_halo(ctx, sa, 1, gp)) || (rc = launch_halo(ctx, isa, 1, gp))) return rc;
        if ((rc = launch_lap(ctx, MGX_LAP_4TH, (const double*)sa, Lv[0].t, 1, gp))) return rc;
        hipLaunchKernelGGL(k_pb_pp, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, (const double*)Lv[0].t,
            (const double*)isa, a, Lv[0].pp);
    }
    else
        hipLaunchKernelGGL(k_pbg_eps, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, rhod, a, e0, rho0, drho0,
            Lv[0].pp);
    MGX_CHECK_LAUNCH(ctx);
    if ((rc = launch_halo(ctx, Lv[0].pp, 1, gp))) return rc;
    // the coarse operators' pp (eps): full-weighting restriction of the
    // finer level's (traded periodically first), with periodic ghosts
    for (int l = 1; l < nl; l++)
    {
        mgx_grid_t cp = Lv[l].g;
        cp.bc[0] = cp.bc[1] = cp.bc[2] = 1;
        if ((rc = launch_restrict(ctx, (const double*)Lv[l - 1].pp, l == 1 ? gp : Lv[l - 1].g, Lv[l].pp, Lv[l].g, 1)))
            return rc;
        if ((rc = launch_halo(ctx, Lv[l].pp, 1, cp))) return rc;
    }
    // Mgm: for PBh4M vh /= inv_sqrt_a, rhs = B(rho inv_sqrt_a); PBh4 and PBh2
    // keep FDoper's identity transforms (src/pb/FDoper.h:66-79): rhs = rho
    if (mehr)
    {
        hipLaunchKernelGGL(k_pb_scale, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, (const double*)vh,
            (const double*)isa, a, 1, vh);
        hipLaunchKernelGGL(k_pb_scale, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, rho, (const double*)isa, a,
            0, r0);
        MGX_CHECK_LAUNCH(ctx);
        if ((rc = launch_halo(ctx, r0, 1, g0))) return rc;
        hipLaunchKernelGGL(k_rhs<double>, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, (const double*)r0, rhs,
            a, a, a.size);
    }
    else
        hipLaunchKernelGGL(k_pb_scale, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, rho, rho, a, 2, rhs);
    MGX_CHECK_LAUNCH(ctx);
    double rhs_norm = 0.;
    if ((rc = int_norm(ctx, P, rhs, g0, &rhs_norm))) return rc;
    MGX_REQUIRE(ctx, rhs_norm > 0., "zero right-hand side");
    const double inv_rhs_norm = 1. / rhs_norm;
    double init = 1., fin = 1., rel = 1.;
    int conv = 0, sweeps = 0;
    double* lhs = Lv[0].t;
    for (int i = 0; i < max_sweeps; i++)
    {
        if ((rc = pb_apply(ctx, Lv[0], vh, lhs))) return rc;
        hipLaunchKernelGGL(k_diff, dim3(nblocks(n0)), dim3(kBlock), 0, ctx->stream, (const double*)rhs,
            (const double*)lhs, r0, n0);
        MGX_CHECK_LAUNCH(ctx);
        double rn = 0.;
        if ((rc = int_norm(ctx, P, r0, g0, &rn))) return rc;
        if (i == 0) init = rn;
        if (rn * inv_rhs_norm < tol)
        {
            fin  = rn;
            rel  = rn * inv_rhs_norm;
            conv = 1;
            break;
        }
        if ((rc = launch_fill(ctx, corr, n0, 0.))) return rc;
        if ((rc = pb_vcycle(ctx, Lv, P, 0, corr, r0, nu1, nu2))) return rc;
        sweeps++;
        if ((rc = launch_axpy(ctx, n0, 1., (const double*)corr, vh))) return rc;
    }
    if (!conv)
    {
        if ((rc = pb_apply(ctx, Lv[0], vh, lhs))) return rc;
        if ((rc = launch_axpy(ctx, n0, -1., (const double*)rhs, lhs))) return rc;
        if ((rc = int_norm(ctx, P, lhs, g0, &fin))) return rc;
        rel = fin * inv_rhs_norm;
    }
    // PBh4M: vh *= inv_sqrt_a; fresh ghosts; get_vepsilon
    if (mehr)
    {
        hipLaunchKernelGGL(k_pb_scale, dim3(nblocks(a.size)), dim3(kBlock), 0, ctx->stream, (const double*)vh,
            (const double*)isa, a, 0, vh);
        MGX_CHECK_LAUNCH(ctx);
    }
    if ((rc = launch_halo(ctx, vh, 1, g0))) return rc;
    if (vks)
    {
        hipLaunchKernelGGL(op == MGX_LAP_2ND ? k_pb_vks<2> : k_pb_vks<4>, dim3(nblocks(a.size)), dim3(kBlock), 0,
            ctx->stream, (const double*)vh, rhod, a, g0.h[0], g0.h[1], g0.h[2], e0, rho0, drho0, vks);
        MGX_CHECK_LAUNCH(ctx);
    }
    if (st)
    {
        st->converged               = conv;
        st->nb_sweeps               = sweeps;
        st->final_residual          = fin;
        st->final_relative_residual = rel;
        st->residual_reduction      = fin / init;
    }
    return MGX_OK;
}

// B psi on packed FP64 data (ld), numpt rows
int rhsM_packed(mgx_ctx* ctx, const mgx_grid_t* grid, const double* psi, int64_t ld, int n, double* out)
{
    PGeom g;
    g.d0   = grid->dim[0];
    g.d1   = grid->dim[1];
    g.d2   = grid->dim[2];
    g.bc0  = grid->bc[0] == 1;
    g.bc1  = grid->bc[1] == 1;
    g.bc2  = grid->bc[2] == 1;
    g.size = (int64_t)g.d0 * g.d1 * g.d2;
    MGX_REQUIRE(ctx, ld >= g.size, "leading dimension smaller than numpt");
    const int64_t total = (int64_t)n * g.size;
    hipLaunchKernelGGL(k_rhsM_packed<double>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, psi, ld, out, ld,
        g, total);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

}  // namespace

int halo(mgx_ctx* ctx, int dtype, void* data, int nfunc, const mgx_grid_t& g)
{
    return dtype == MGX_F32 ? launch_halo(ctx, (float*)data, nfunc, g) : launch_halo(ctx, (double*)data, nfunc, g);
}

}  // namespace mgx

using namespace mgx;

namespace mgx
{
namespace
{
// separable orbitals on z-planes [z0, z0 + nzs) of a periodic grid and
// (-Delta_h4 + V) of them: the 4th-order stencil is a sum of 1D stencils, so
// H psi = (L fx) fy fz + fx (L fy) fz + fx fy (L fz) + V psi exactly, with
// the 1D second differences L f of the factors given
__global__ void k_sep_orbitals(int n, int gx, int gy, int gz, int z0, int nzs, const double* fx, const double* fy,
    const double* fz, const double* lfx, const double* lfy, const double* lfz, const double* vtot, double* psi,
    double* hpsi)
{
    const int64_t K   = (int64_t)gx * gy * nzs;
    const int64_t tot = K * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int j     = (int)(t / K);
        const int64_t k = t - (int64_t)j * K;
        const int zl = (int)(k % nzs), y = (int)((k / nzs) % gy), x = (int)(k / ((int64_t)nzs * gy));
        const int z  = z0 + zl;
        const double a = fx[(int64_t)j * gx + x], b = fy[(int64_t)j * gy + y], c = fz[(int64_t)j * gz + z];
        const double p = a * b * c;
        psi[t]         = p;
        hpsi[t] = lfx[(int64_t)j * gx + x] * b * c + a * lfy[(int64_t)j * gy + y] * c + a * b * lfz[(int64_t)j * gz + z]
                  + vtot[((int64_t)x * gy + y) * gz + z] * p;
    }
}
}  // namespace
}  // namespace mgx

extern "C" {

int mgx_halo(mgx_ctx* ctx, mgx_block_t* b)
{
    int rc = check_block(ctx, b);
    if (rc) return rc;
    rc = halo(ctx, b->dtype, b->data, b->nfunc, b->grid);
    return rc ? rc : sync(ctx);
}

int mgx_lap_apply(mgx_ctx* ctx, int type, const mgx_block_t* in, mgx_block_t* out)
{
    int rc = check_block(ctx, in);
    if (!rc) rc = check_block(ctx, out);
    if (rc) return rc;
    MGX_REQUIRE(ctx, same_shape(in, out), "in/out blocks differ in shape");
    rc = in->dtype == MGX_F32
             ? launch_lap(ctx, type, (const float*)in->data, (float*)out->data, in->nfunc, in->grid)
             : launch_lap(ctx, type, (const double*)in->data, (double*)out->data, in->nfunc, in->grid);
    return rc ? rc : sync(ctx);
}

int mgx_rhs_apply(mgx_ctx* ctx, const mgx_block_t* in, mgx_block_t* out)
{
    int rc = check_block(ctx, in);
    if (!rc) rc = check_block(ctx, out);
    if (rc) return rc;
    MGX_REQUIRE(ctx, in->grid.ghosts >= 1, "Mehrstellen rhs needs 1 ghost");
    MGX_REQUIRE(ctx, in->dtype == out->dtype && in->nfunc == out->nfunc, "in/out mismatch");
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, in->grid.dim[i] == out->grid.dim[i], "in/out dims differ");
    GGeom a = ggeom(in->grid), b = ggeom(out->grid);
    const int64_t total = (int64_t)in->nfunc * a.size;
    if (in->dtype == MGX_F32)
        hipLaunchKernelGGL(k_rhs<float>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream,
            (const float*)in->data, (float*)out->data, a, b, total);
    else
        hipLaunchKernelGGL(k_rhs<double>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream,
            (const double*)in->data, (double*)out->data, a, b, total);
    MGX_CHECK_LAUNCH(ctx);
    return sync(ctx);
}

int mgx_pointwise_product(mgx_ctx* ctx, const mgx_block_t* a, const double* v, mgx_block_t* out)
{
    int rc = check_block(ctx, a);
    if (!rc) rc = check_block(ctx, out);
    if (rc) return rc;
    MGX_REQUIRE(ctx, v != nullptr, "null potential");
    MGX_REQUIRE(ctx, same_shape(a, out), "in/out blocks differ in shape");
    GGeom g             = ggeom(a->grid);
    const int64_t total = (int64_t)a->nfunc * g.sizeg;
    if (a->dtype == MGX_F32)
        hipLaunchKernelGGL(k_pointwise<float>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream,
            (const float*)a->data, v, (float*)out->data, g.sizeg, total);
    else
        hipLaunchKernelGGL(k_pointwise<double>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream,
            (const double*)a->data, v, (double*)out->data, g.sizeg, total);
    MGX_CHECK_LAUNCH(ctx);
    return sync(ctx);
}

int mgx_synth_separable(mgx_ctx* ctx, int n, const int gdim[3], int z0, int nzs, const double* fx,
    const double* fy, const double* fz, const double* lfx, const double* lfy, const double* lfz, const double* vtot,
    double* psi, double* hpsi)
{
    MGX_REQUIRE(ctx, n > 0 && gdim && gdim[0] > 0 && gdim[1] > 0 && gdim[2] > 0, "bad separable orbital sizes");
    MGX_REQUIRE(ctx, z0 >= 0 && nzs > 0 && z0 + nzs <= gdim[2], "z-planes outside the grid");
    MGX_REQUIRE(ctx, fx && fy && fz && lfx && lfy && lfz && vtot && psi && hpsi, "null data");
    const int64_t tot = (int64_t)gdim[0] * gdim[1] * nzs * n;
    hipLaunchKernelGGL(k_sep_orbitals, dim3(nblocks(tot)), dim3(kBlock), 0, ctx->stream, n, gdim[0], gdim[1],
        gdim[2], z0, nzs, fx, fy, fz, lfx, lfy, lfz, vtot, psi, hpsi);
    MGX_CHECK_LAUNCH(ctx);
    return sync(ctx);
}

int mgx_axpy(mgx_ctx* ctx, int dtype, int64_t n, double alpha, const void* x, void* y)
{
    MGX_REQUIRE(ctx, n >= 0 && (n == 0 || (x && y)), "bad axpy arguments");
    if (n == 0) return MGX_OK;
    int rc = dtype == MGX_F32 ? launch_axpy(ctx, n, alpha, (const float*)x, (float*)y)
                              : launch_axpy(ctx, n, alpha, (const double*)x, (double*)y);
    return rc ? rc : sync(ctx);
}

int mgx_stream_triad(mgx_ctx* ctx, int64_t n, float* a, const float* b, const float* c, int reps, double* gbps,
    int* unroll)
{
    MGX_REQUIRE(ctx, n > 0 && n % 4 == 0 && a && b && c && gbps && reps > 0, "bad stream triad arguments");
    MGX_REQUIRE(ctx, ((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) % 16 == 0, "stream triad needs 16-byte alignment");
    const int64_t n4 = n / 4;
    hipEvent_t e0, e1;
    MGX_CHECK_HIP(ctx, hipEventCreate(&e0));
    MGX_CHECK_HIP(ctx, hipEventCreate(&e1));
    double best = 0.;
    int best_u  = 0;
    const int us[3] = { 2, 4, 8 };
    for (int k = 0; k < 3; k++)
    {
        const int u           = us[k];
        const unsigned blocks = (unsigned)((n4 + 256 * u - 1) / (256 * u));
        auto launch           = [&]() {
            if (u == 2)
                hipLaunchKernelGGL(k_triad4<2>, dim3(blocks), dim3(256), 0, ctx->stream, (const float4*)b,
                    (const float4*)c, (float4*)a, n4, 0.5f);
            else if (u == 4)
                hipLaunchKernelGGL(k_triad4<4>, dim3(blocks), dim3(256), 0, ctx->stream, (const float4*)b,
                    (const float4*)c, (float4*)a, n4, 0.5f);
            else
                hipLaunchKernelGGL(k_triad4<8>, dim3(blocks), dim3(256), 0, ctx->stream, (const float4*)b,
                    (const float4*)c, (float4*)a, n4, 0.5f);
        };
        launch();
        launch();
        MGX_CHECK_HIP(ctx, hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; r++)
            launch();
        MGX_CHECK_HIP(ctx, hipEventRecord(e1, ctx->stream));
        MGX_CHECK_LAUNCH(ctx);
        MGX_CHECK_HIP(ctx, hipEventSynchronize(e1));
        float ms = 0.f;
        MGX_CHECK_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        const double rate = 12. * (double)n * reps / (ms * 1e-3) / 1e9;
        if (rate > best)
        {
            best   = rate;
            best_u = u;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *gbps = best;
    if (unroll) *unroll = best_u;
    return MGX_OK;
}

int mgx_jacobi(mgx_ctx* ctx, int type, double jf, const mgx_block_t* f, mgx_block_t* x, mgx_block_t* w)
{
    int rc = check_block(ctx, f);
    if (!rc) rc = check_block(ctx, x);
    if (!rc) rc = check_block(ctx, w);
    if (rc) return rc;
    MGX_REQUIRE(ctx, same_shape(f, x) && same_shape(x, w), "jacobi blocks differ in shape");
    rc = x->dtype == MGX_F32 ? jacobi_entry(ctx, type, jf, (const float*)f->data, (float*)x->data, (float*)w->data,
                                   x->nfunc, x->grid)
                             : jacobi_entry(ctx, type, jf, (const double*)f->data, (double*)x->data,
                                   (double*)w->data, x->nfunc, x->grid);
    return rc ? rc : sync(ctx);
}

int mgx_restrict(mgx_ctx* ctx, const mgx_block_t* fine, mgx_block_t* coarse)
{
    int rc = check_block(ctx, fine);
    if (!rc) rc = check_block(ctx, coarse);
    if (rc) return rc;
    MGX_REQUIRE(ctx, fine->dtype == coarse->dtype && fine->nfunc == coarse->nfunc, "fine/coarse mismatch");
    rc = fine->dtype == MGX_F32
             ? launch_restrict(ctx, (const float*)fine->data, fine->grid, (float*)coarse->data, coarse->grid,
                   fine->nfunc)
             : launch_restrict(ctx, (const double*)fine->data, fine->grid, (double*)coarse->data, coarse->grid,
                   fine->nfunc);
    return rc ? rc : sync(ctx);
}

int mgx_prolong(mgx_ctx* ctx, const mgx_block_t* coarse, mgx_block_t* fine)
{
    int rc = check_block(ctx, fine);
    if (!rc) rc = check_block(ctx, coarse);
    if (rc) return rc;
    MGX_REQUIRE(ctx, fine->dtype == coarse->dtype && fine->nfunc == coarse->nfunc, "fine/coarse mismatch");
    rc = fine->dtype == MGX_F32
             ? launch_extend(ctx, (const float*)coarse->data, coarse->grid, (float*)fine->data, fine->grid,
                   fine->nfunc)
             : launch_extend(ctx, (const double*)coarse->data, coarse->grid, (double*)fine->data, fine->grid,
                   fine->nfunc);
    return rc ? rc : sync(ctx);
}

int mgx_mask(mgx_ctx* ctx, mgx_block_t* u, const uint8_t* mask)
{
    int rc = check_block(ctx, u);
    if (rc) return rc;
    rc = u->dtype == MGX_F32 ? launch_mask(ctx, (float*)u->data, mask, u->nfunc, u->grid)
                             : launch_mask(ctx, (double*)u->data, mask, u->nfunc, u->grid);
    return rc ? rc : sync(ctx);
}

int mgx_mg(mgx_ctx* ctx, int type, int max_levels, mgx_block_t* v, const mgx_block_t* f,
    const uint8_t* const* masks)
{
    int rc = check_block(ctx, v);
    if (!rc) rc = check_block(ctx, f);
    if (rc) return rc;
    MGX_REQUIRE(ctx, same_shape(v, f), "v/f blocks differ in shape");
    rc = v->dtype == MGX_F32
             ? mg_run(ctx, type, max_levels, (float*)v->data, (const float*)f->data, v->nfunc, v->grid, masks, 1)
             : mg_run(ctx, type, max_levels, (double*)v->data, (const double*)f->data, v->nfunc, v->grid, masks,
                   1);
    return rc ? rc : sync(ctx);
}

int mgx_poisson_mg(mgx_ctx* ctx, int lap_type, const mgx_block_t* rho, mgx_block_t* vh, int nu1, int nu2,
    int max_sweeps, double tol, int max_nlevels, mgx_poisson_stats_t* stats)
{
    int rc = check_block(ctx, rho);
    if (!rc) rc = check_block(ctx, vh);
    if (rc) return rc;
    MGX_REQUIRE(ctx, same_shape(rho, vh), "rho/vh blocks differ in shape");
    MGX_REQUIRE(ctx, rho->dtype == MGX_F64 && rho->nfunc == 1, "one FP64 function");
    MGX_REQUIRE(ctx, lap_type >= 0 && lap_type <= 2, "LapFactory::createLap() --- option invalid");
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, vh->grid.bc[i] == 0 || vh->grid.bc[i] == 1,
            "boundary conditions are 0 (Dirichlet, zero) or 1 (periodic); multipole values are not supported");
    MGX_REQUIRE(ctx, vh->grid.ghosts >= (lap_type == MGX_LAP_4TH ? 2 : 1), "too few ghosts for the operator");
    MGX_REQUIRE(ctx, nu1 >= 0 && nu2 >= 0 && max_sweeps >= 0 && max_nlevels >= 0, "bad cycle parameters");
    rc = poisson_run(ctx, lap_type, vh->grid, (const double*)rho->data, (double*)vh->data, nu1, nu2, max_sweeps, tol,
        max_nlevels, stats);
    return rc ? rc : sync(ctx);
}

int mgx_pb_mg(mgx_ctx* ctx, int lap_type, const mgx_block_t* rho, const mgx_block_t* rhod, mgx_block_t* vh,
    mgx_block_t* vks, double e0, double rho0, double drho0, int nu1, int nu2, int max_sweeps, double tol,
    int max_nlevels, mgx_poisson_stats_t* stats)
{
    int rc = check_block(ctx, rho);
    if (!rc) rc = check_block(ctx, rhod);
    if (!rc) rc = check_block(ctx, vh);
    if (!rc && vks) rc = check_block(ctx, vks);
    if (rc) return rc;
    MGX_REQUIRE(ctx, lap_type == MGX_LAP_MEHRSTELLEN || lap_type == MGX_LAP_4TH || lap_type == MGX_LAP_2ND,
        "Electrostatic, Undefined option (the dielectric operators are PBh4M, PBh4 and PBh2)");
    MGX_REQUIRE(ctx, same_shape(rho, vh) && same_shape(rhod, vh) && (!vks || same_shape(vks, vh)),
        "rho/rhod/vh/vks blocks differ in shape");
    MGX_REQUIRE(ctx, vh->dtype == MGX_F64 && vh->nfunc == 1, "one FP64 function");
    MGX_REQUIRE(ctx, vh->grid.ghosts >= (lap_type == MGX_LAP_2ND ? 1 : 2),
        "Not enough ghosts points in the dielectric operator (2 for PBh4M and PBh4, 1 for PBh2)");
    MGX_REQUIRE(ctx, e0 >= 1. && rho0 > 0. && drho0 > 0., "bad dielectric parameters");
    MGX_REQUIRE(ctx, nu1 >= 0 && nu2 >= 0 && max_sweeps >= 0 && max_nlevels >= 0, "bad cycle parameters");
    rc = pb_run(ctx, lap_type, vh->grid, (const double*)rho->data, (const double*)rhod->data, (double*)vh->data,
        vks ? (double*)vks->data : nullptr, e0, rho0, drho0, nu1, nu2, max_sweeps, tol, max_nlevels, stats);
    return rc ? rc : sync(ctx);
}

int mgx_hloc_apply(mgx_ctx* ctx, int type, const mgx_grid_t* grid, int dtype, const void* psi, int64_t ld,
    int nfunc, const double* vtot, void* hpsi, int64_t ldo)
{
    MGX_REQUIRE(ctx, grid && psi && vtot && hpsi && nfunc > 0, "bad hloc arguments");
    MGX_REQUIRE(ctx, dtype == MGX_F32 || dtype == MGX_F64, "bad dtype");
    if (int rc = check_bc(ctx, *grid)) return rc;
    PGeom g;
    g.d0   = grid->dim[0];
    g.d1   = grid->dim[1];
    g.d2   = grid->dim[2];
    g.bc0  = grid->bc[0] == 1;
    g.bc1  = grid->bc[1] == 1;
    g.bc2  = grid->bc[2] == 1;
    g.size = (int64_t)g.d0 * g.d1 * g.d2;
    MGX_REQUIRE(ctx, ld >= g.size && ldo >= g.size, "leading dimension smaller than numpt");
    for (int i = 0; i < 3; i++)
        MGX_REQUIRE(ctx, grid->dim[i] >= 2, "grid too small");
    const int64_t total = (int64_t)nfunc * g.size;
    dim3 grd(nblocks(total)), blk(kBlock);
    if (type == MGX_LAP_4TH)
    {
        // the marching kernel (packed rows, halo by wrapping) when it applies
        const MView vi = mview_packed(*grid, ld), vo = mview_packed(*grid, ldo);
        const int m    = dtype == MGX_F32
                             ? march_grid(ctx, MGX_LAP_4TH, MG_HLOC, (const float*)psi, vi, (const float*)nullptr, vtot,
                                   (float*)hpsi, vo, (float*)nullptr, nfunc, *grid, 0.)
                             : march_grid(ctx, MGX_LAP_4TH, MG_HLOC, (const double*)psi, vi, (const double*)nullptr,
                                   vtot, (double*)hpsi, vo, (double*)nullptr, nfunc, *grid, 0.);
        if (m < 0) return m;
        if (m) return sync(ctx);
        Coef4 c = coef4(grid->h);
        if (dtype == MGX_F32)
            hipLaunchKernelGGL(k_hloc4_packed<float>, grd, blk, 0, ctx->stream, (const float*)psi, ld, vtot,
                (float*)hpsi, ldo, g, c, total);
        else
            hipLaunchKernelGGL(k_hloc4_packed<double>, grd, blk, 0, ctx->stream, (const double*)psi, ld, vtot,
                (double*)hpsi, ldo, g, c, total);
    }
    else if (type == MGX_LAP_MEHRSTELLEN)
    {
        CoefM c = coefm(grid->h);
        if (dtype == MGX_F32)
            hipLaunchKernelGGL(k_hlocM_packed<float>, grd, blk, 0, ctx->stream, (const float*)psi, ld, vtot,
                (float*)hpsi, ldo, g, c, total);
        else
            hipLaunchKernelGGL(k_hlocM_packed<double>, grd, blk, 0, ctx->stream, (const double*)psi, ld, vtot,
                (double*)hpsi, ldo, g, c, total);
    }
    else
        return set_err(ctx, MGX_EINVAL, "Hamiltonian::applyLocal supports lap types 0 and 2");
    MGX_CHECK_LAUNCH(ctx);
    return sync(ctx);
}

int mgx_mehr_bmatrix(mgx_ctx* ctx, const mgx_grid_t* grid, const double* psi, int64_t ld, int n, double* B)
{
    MGX_REQUIRE(ctx, grid && psi && B && n > 0, "bad B-matrix arguments");
    const int64_t numpt = (int64_t)grid->dim[0] * grid->dim[1] * grid->dim[2];
    MGX_REQUIRE(ctx, ld >= numpt, "leading dimension smaller than numpt");
    double* bpsi = (double*)workspace(ctx, 0, (size_t)ld * n * sizeof(double));
    if (!bpsi) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    int rc = rhsM_packed(ctx, grid, psi, ld, n, bpsi);
    if (rc) return rc;
    const double vel = grid->h[0] * grid->h[1] * grid->h[2];
    return mgx_local_product(ctx, MGX_PREC_F64, psi, bpsi, ld, numpt, n, 1, vel, B);
}

int mgx_mehr_residual(mgx_ctx* ctx, const mgx_grid_t* grid, const double* psi, const double* hpsi,
    const double* theta, int64_t ld, int n, double* res)
{
    MGX_REQUIRE(ctx, grid && psi && hpsi && theta && res && n > 0, "bad residual arguments");
    const int64_t numpt = (int64_t)grid->dim[0] * grid->dim[1] * grid->dim[2];
    MGX_REQUIRE(ctx, ld >= numpt, "leading dimension smaller than numpt");
    double* bpsi = (double*)workspace(ctx, 0, (size_t)ld * n * sizeof(double));
    if (!bpsi) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    int rc = rhsM_packed(ctx, grid, psi, ld, n, bpsi);
    if (rc) return rc;
    return mgx_residual(ctx, bpsi, hpsi, theta, ld, numpt, n, res);
}

int mgx_precond_mg(mgx_ctx* ctx, int type, int max_levels, double gamma, int prec_dtype, const mgx_grid_t* grid,
    const double* res, int64_t ld, int nfunc, const uint8_t* const* masks, double* out, int64_t ldo)
{
    MGX_REQUIRE(ctx, grid && res && out && nfunc > 0, "bad precond arguments");
    MGX_REQUIRE(ctx, gamma > 0., "gamma must be positive");
    if (int r = check_bc(ctx, *grid)) return r;
    const int64_t numpt = (int64_t)grid->dim[0] * grid->dim[1] * grid->dim[2];
    MGX_REQUIRE(ctx, ld >= numpt && ldo >= numpt, "leading dimension smaller than numpt");
    int rc = prec_dtype == MGX_F32
                 ? precond_run<float>(ctx, type, max_levels, gamma, *grid, res, ld, nfunc, masks, out, ldo)
                 : precond_run<double>(ctx, type, max_levels, gamma, *grid, res, ld, nfunc, masks, out, ldo);
    return rc ? rc : sync(ctx);
}

}  // extern "C"
