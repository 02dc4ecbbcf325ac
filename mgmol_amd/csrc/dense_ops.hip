// dense_ops.hip -- tall-skinny contractions on MFMA and the FP64 dense layer.
//
// ss(i,j) = vel * sum_k A(k,i) B(k,j) over one slab of rows (the reference's
// per-subdomain LocalMatrices, src/LocGridOrbitals.cc:1504-1604).  Each wave
// owns one 16x16 output tile and a chunk of rows:
//   PREC_F64  : v_mfma_f64_16x16x4_f64, FP64 accumulation;
//   PREC_SPLIT: x = hi + lo (hi = fl32(x), lo = fl32(x - hi)),
//               v_mfma_f32_16x16x4_f32 on hi*hi and hi*lo + lo*hi, FP32
//               accumulation over 64-row sub-chunks, FP64 sum of sub-chunks.
// Chunk partials are reduced in a fixed order (deterministic) and scaled.
#include "common.hpp"

#include <rocsolver/rocsolver.h>

#include <mutex>

namespace mgx
{
namespace
{

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;  // waves per block, each on its own row sub-chunk

// MFMA 16x16x4 operand layout (CDNA3/CDNA4): lane l supplies A[m=l%16][k=l/16]
// and B[k=l/16][n=l%16].  Accumulator v of lane l holds D[4*(l/16)+v][l%16]
// for v_mfma_f32_16x16x4_f32 and D[l/16+4*v][l%16] for v_mfma_f64_16x16x4_f64
// (both checked on gfx950 by tests/test_gpu_dense.py).
template <int PREC>
__global__ void __launch_bounds__(64 * kWaves) k_tsgemm(const double* A, const double* B, int64_t ld,
    int64_t loc_numpt, int n, int ntile, int nchunk, int64_t rows_per_chunk, int lower_only, double* part)
{
    const int tile  = blockIdx.x;  // tile index over (ti, tj)
    const int chunk = blockIdx.y;
    const int slab  = blockIdx.z;
    const int ti = tile / ntile, tj = tile - (tile / ntile) * ntile;
    if (lower_only && tj > ti) return;
    const int lane = threadIdx.x & 63;
    const int wid  = threadIdx.x >> 6;
    const int c = lane & 15, r = lane >> 4;
    const int i0 = ti * 16, j0 = tj * 16;
    const int64_t k_begin = (int64_t)slab * loc_numpt + (int64_t)chunk * rows_per_chunk;
    int64_t k_end         = k_begin + rows_per_chunk;
    const int64_t slab_end = (int64_t)(slab + 1) * loc_numpt;
    if (k_end > slab_end) k_end = slab_end;
    const int ia = i0 + c, jb = j0 + c;
    const bool va = ia < n, vb = jb < n;
    const double* pa = A + (int64_t)(va ? ia : 0) * ld;
    const double* pb = B + (int64_t)(vb ? jb : 0) * ld;
    // the wave's rows: interleave waves in blocks of 64 rows
    double acc[4] = { 0., 0., 0., 0. };
    if (PREC == MGX_PREC_F64)
    {
        d4 d = { 0., 0., 0., 0. };
        for (int64_t k0 = k_begin + 64 * wid; k0 < k_end; k0 += 64 * kWaves)
        {
#pragma unroll 4
            for (int s = 0; s < 16; s++)
            {
                const int64_t k = k0 + 4 * s + r;
                const bool in   = k < k_end;
                const double a  = (va && in) ? pa[k] : 0.;
                const double b  = (vb && in) ? pb[k] : 0.;
                d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
            }
        }
        for (int v = 0; v < 4; v++)
            acc[v] = d[v];
    }
    else
    {
        for (int64_t k0 = k_begin + 64 * wid; k0 < k_end; k0 += 64 * kWaves)
        {
            f4 hh = { 0.f, 0.f, 0.f, 0.f }, hx = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll 4
            for (int s = 0; s < 16; s++)
            {
                const int64_t k = k0 + 4 * s + r;
                const bool in   = k < k_end;
                const double a  = (va && in) ? pa[k] : 0.;
                const double b  = (vb && in) ? pb[k] : 0.;
                const float ah = (float)a, al = (float)(a - (double)ah);
                const float bh = (float)b, bl = (float)(b - (double)bh);
                hh = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bh, hh, 0, 0, 0);
                hx = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bl, hx, 0, 0, 0);
                hx = __builtin_amdgcn_mfma_f32_16x16x4f32(al, bh, hx, 0, 0, 0);
            }
            for (int v = 0; v < 4; v++)
                acc[v] += (double)hh[v] + (double)hx[v];
        }
    }
    // combine the waves' partial tiles in LDS (fixed order)
    __shared__ double red[kWaves][64][4];
    for (int v = 0; v < 4; v++)
        red[wid][lane][v] = acc[v];
    __syncthreads();
    if (wid == 0)
    {
        for (int v = 0; v < 4; v++)
        {
            double s = red[0][lane][v];
            for (int w = 1; w < kWaves; w++)
                s += red[w][lane][v];
            const int row = (PREC == MGX_PREC_F64) ? i0 + r + 4 * v : i0 + 4 * r + v;
            const int col = j0 + c;
            if (row < n && col < n)
                part[(((int64_t)slab * nchunk + chunk) * n + col) * n + row] = s;
        }
    }
}

// ss[slab](i,j) = vel * sum_chunks part; lower_only fills the upper triangle
// from the lower one (SquareLocalMatrices::fillUpperWithLower)
__global__ void k_chunk_reduce(const double* part, int n, int nchunk, int subdivx, double vel, int lower_only,
    double* ss)
{
    const int64_t nn = (int64_t)n * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn * subdivx;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t slab = t / nn;
        const int64_t e    = t - slab * nn;
        int col = (int)(e / n), row = (int)(e - (e / n) * n);
        int64_t src = e;
        if (lower_only && col > row) src = (int64_t)row * n + col;  // (col,row) element in col-major
        double s = 0.;
        for (int c = 0; c < nchunk; c++)
            s += part[(slab * nchunk + c) * nn + src];
        ss[slab * nn + e] = s * vel;
    }
}

__global__ void k_l2g(const double* ss, const int* gids, int subdivx, int m, int numst, double tol, double* glob)
{
    // LocalMatrices2DistMatrix::convert, src/LocalMatrices2DistMatrix.cc:21-78;
    // one thread per (icolor, jcolor), slabs in order, consecutive same-pair
    // entries summed before the push.
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)m * m;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int ic = (int)(t / m), jc = (int)(t - (t / m) * m);
        int pst_old = -1;
        double val  = 0.;
        int st1o = -1, st2o = -1;
        for (int s = 0; s < subdivx; s++)
        {
            const int st1 = gids[s * m + ic];
            if (st1 == -1) continue;
            const int st2 = gids[s * m + jc];
            if (st2 == -1) continue;
            const int pst = st2 * numst + st1;
            const double v = ss[((int64_t)s * m + jc) * m + ic];
            if (fabs(v) > tol)
            {
                if (pst == pst_old)
                    val += v;
                else
                {
                    if (pst_old >= 0) atomicAdd(&glob[(int64_t)st2o * numst + st1o], val);
                    val     = v;
                    st1o    = st1;
                    st2o    = st2;
                    pst_old = pst;
                }
            }
        }
        if (pst_old >= 0) atomicAdd(&glob[(int64_t)st2o * numst + st1o], val);
    }
}

__global__ void k_tri(double* L, int n, int zero_upper)
{
    // zero_upper: L(i,j) = 0 for j > i ; else symmetrise from the lower triangle
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n * n;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int col = (int)(t / n), row = (int)(t - (t / n) * n);
        if (col <= row) continue;
        L[t] = zero_upper ? 0. : L[(int64_t)row * n + col];
    }
}

// single block: out = sum_{i,k} H(i,k) * Sinv(k,i)
__global__ void __launch_bounds__(1024) k_trace_prod(const double* H, const double* Sinv, int n, double* out)
{
    __shared__ double red[1024];
    double s = 0.;
    for (int64_t t = threadIdx.x; t < (int64_t)n * n; t += blockDim.x)
    {
        const int k = (int)(t / n), i = (int)(t - (t / n) * n);  // H(i,k) at t
        s += H[t] * Sinv[(int64_t)i * n + k];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// single block: out = sum_i A(i,i)
__global__ void __launch_bounds__(256) k_diag_sum(const double* A, int n, double* out)
{
    __shared__ double red[256];
    double s = 0.;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        s += A[(int64_t)i * n + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// per-block max |delta_ij - T(i,j)|, then a second pass over the blocks
__global__ void __launch_bounds__(256) k_id_err(const double* T, int n, double* blockmax)
{
    __shared__ double red[256];
    double m = 0.;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n * n;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int col = (int)(t / n), row = (int)(t - (t / n) * n);
        const double v = fabs(T[t] - (row == col ? 1. : 0.));
        m = v > m ? v : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) blockmax[blockIdx.x] = red[0];
}

// k_id_err that also writes E = I - T and X in FP32 (the operands of the
// FP32 correction product of a Newton-Schulz step)
__global__ void __launch_bounds__(256) k_id_err_f32(const double* T, const double* X, int n, double* blockmax,
    float* E32, float* X32)
{
    __shared__ double red[256];
    double m = 0.;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n * n;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int col = (int)(t / n), row = (int)(t - (t / n) * n);
        const double e = (row == col ? 1. : 0.) - T[t];
        const double v = fabs(e);
        m              = v > m ? v : m;
        E32[t]         = (float)e;
        X32[t]         = (float)X[t];
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) blockmax[blockIdx.x] = red[0];
}

// round to the nearest BF16 (ties to even) through FP32
__device__ __forceinline__ uint16_t bf16_rn(double v)
{
    const float f    = (float)v;
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ double bf16_val(uint16_t h) { return (double)__uint_as_float((uint32_t)h << 16); }

// k_id_err with E = I - T and X as BF16 hi/lo pairs (hi = bf16(v),
// lo = bf16(v - hi)), laid out as the operands of one product with K = 3n:
// A = [Eh Eh El] (n x 3n) and B = [Xh; Xl; Xh] (3n x n), both column-major,
// so that A B = Eh Xh + Eh Xl + El Xh
__global__ void __launch_bounds__(256) k_id_err_bf16(const double* T, const double* X, int n, double* blockmax,
    uint16_t* A, uint16_t* B)
{
    __shared__ double red[256];
    double m = 0.;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n * n;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int col = (int)(t / n), row = (int)(t - (t / n) * n);
        const double e = (row == col ? 1. : 0.) - T[t];
        const double v = fabs(e);
        m              = v > m ? v : m;
        const int64_t nn  = (int64_t)n * n;
        const uint16_t eh = bf16_rn(e);
        A[t]              = eh;
        A[nn + t]         = eh;
        A[2 * nn + t]     = bf16_rn(e - bf16_val(eh));
        const double x    = X[t];
        const uint16_t xh = bf16_rn(x);
        const int64_t bj  = (int64_t)col * 3 * n + row;
        B[bj]             = xh;
        B[bj + n]         = bf16_rn(x - bf16_val(xh));
        B[bj + 2 * n]     = xh;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) blockmax[blockIdx.x] = red[0];
}

// out = x + (double)y
__global__ void k_add_f32(const double* x, const float* y, double* out, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = x[t] + (double)y[t];
}

__global__ void __launch_bounds__(256) k_max_reduce(const double* v, int n, double* out)
{
    __shared__ double red[256];
    double m = 0.;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        m = fmax(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// X = (X + X^T)/2
// X = (X + X^T)/2 by 32 x 32 tile pairs through LDS (coalesced on both
// sides); each pair of entries gets 0.5 (x_pq + x_qp), as k_symmetrize
__global__ void __launch_bounds__(256) k_symmetrize_tiles(double* X, int n)
{
    __shared__ double a[32][33], b[32][33];
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj < ti) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8)
    {
        // a[k][tx] = X(ti*32 + tx, tj*32 + k), b[k][tx] = X(tj*32 + tx, ti*32 + k)
        const int r = ti * 32 + tx, c = tj * 32 + k, r2 = tj * 32 + tx, c2 = ti * 32 + k;
        a[k][tx] = (r < n && c < n) ? X[r + (int64_t)c * n] : 0.;
        b[k][tx] = (r2 < n && c2 < n) ? X[r2 + (int64_t)c2 * n] : 0.;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8)
    {
        const int r = ti * 32 + tx, c = tj * 32 + k, r2 = tj * 32 + tx, c2 = ti * 32 + k;
        if (r < n && c < n) X[r + (int64_t)c * n] = 0.5 * (a[k][tx] + b[tx][k]);
        if (r2 < n && c2 < n) X[r2 + (int64_t)c2 * n] = 0.5 * (b[k][tx] + a[tx][k]);
    }
}

__global__ void k_symmetrize(double* X, int n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n * n;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int col = (int)(t / n), row = (int)(t - (t / n) * n);
        if (col >= row) continue;
        const int64_t u = (int64_t)row * n + col;
        const double a = 0.5 * (X[t] + X[u]);
        X[t] = a;
        X[u] = a;
    }
}

__global__ void k_sub(double* res, const double* hpsi, int64_t ld, int64_t numpt, int n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < numpt * n;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t f = t / numpt, p = t - f * numpt;
        res[f * ld + p] += -1. * hpsi[f * ld + p];
    }
}

__global__ void __launch_bounds__(256) k_col_sumsq(const double* r, int64_t ld, int64_t numpt, double vel, double* d)
{
    __shared__ double red[256];
    const int f = blockIdx.x;
    double s    = 0.;
    for (int64_t p = threadIdx.x; p < numpt; p += blockDim.x)
        s += r[f * ld + p] * r[f * ld + p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) d[f] = red[0] * vel;
}

__global__ void __launch_bounds__(256) k_diag_dot(const double* d, const double* Sinv, int n, double* out)
{
    __shared__ double red[256];
    double s = 0.;
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        s += d[j] * Sinv[(int64_t)j * n + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = 0.5 * sqrt(red[0]);
}

// Y = X A, A pair-sparse symmetric, by column tiles: block (t, chunk) holds
// kSpTile output columns of tile t for 256 consecutive rows c, walks the
// union of the tile's rows k once (X[c, k] = X[k, c] read down column k,
// coalesced over c) and applies the tile's coefficients from LDS.
// One block per (column tile, 256-row chunk), mapped XCD-aware so that one
// XCD takes all the tiles of a row chunk: the union columns of the tiles it
// runs together then share that XCD's L2 (a row chunk of all of X is n x 2 KB)
__global__ void __launch_bounds__(256) k_spmm_tiles(const double* __restrict__ X, int n,
    const double* __restrict__ val, const int* tptr, const int* tcol, const int* tk, const int* te,
    double* __restrict__ Y, int ntile, int64_t nreal)
{
    constexpr int U = 64;
    __shared__ double cf[U][kSpTile];
    __shared__ int kk[U];
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    if (L >= nreal) return;
    const int t = (int)(L % ntile);
    const int c = (int)(L / ntile) * 256 + threadIdx.x;
    double acc[kSpTile];
    for (int q = 0; q < kSpTile; q++)
        acc[q] = 0.;
    const int u0 = tptr[t], u1 = tptr[t + 1];
    for (int ub = u0; ub < u1; ub += U)
    {
        const int m = min(U, u1 - ub);
        for (int z = threadIdx.x; z < U * kSpTile; z += 256)
        {
            const int u = z / kSpTile, q = z - (z / kSpTile) * kSpTile;
            double v    = 0.;
            if (u < m)
            {
                const int e = te[(int64_t)(ub + u) * kSpTile + q];
                if (e >= 0) v = val[e];
            }
            cf[u][q] = v;
        }
        if (threadIdx.x < U) kk[threadIdx.x] = threadIdx.x < m ? tk[ub + threadIdx.x] : 0;
        __syncthreads();
        if (c < n)
            for (int u = 0; u < m; u++)
            {
                const double x = X[c + (int64_t)kk[u] * n];
                for (int q = 0; q < kSpTile; q++)
                    acc[q] += cf[u][q] * x;
            }
        __syncthreads();
    }
    if (c < n)
        for (int q = 0; q < kSpTile; q++)
        {
            const int r = tcol[t * kSpTile + q];
            if (r >= 0) Y[c + (int64_t)r * n] = acc[q];
        }
}

// out = 0.5 (a + a^T)  (MatrixType::transpose(0.5, work, 0.5) after symm)
__global__ void k_sym_half(const double* a, int n, double* out)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n * n;
         t += (int64_t)gridDim.x * blockDim.x)
    {
        const int col = (int)(t / n), row = (int)(t - (int64_t)col * n);
        out[t] = 0.5 * a[t] + 0.5 * a[(int64_t)row * n + col];
    }
}

}  // namespace

int spmm_pairs(mgx_ctx* ctx, int n, const double* X, const double* val, const PairTiles& pt, double* Y)
{
    if (pt.ntile <= 0) return set_err(ctx, MGX_EINVAL, "no column tiles");
    // X read and Y written once: 16 n^2 bytes
    const int tk = prof_begin(ctx, MGX_PROF_SPMM, 16. * (double)n * n);
    const int64_t nreal = (int64_t)pt.ntile * ((n + 255) / 256);
    hipLaunchKernelGGL(k_spmm_tiles, dim3((unsigned)((nreal + 7) / 8 * 8)), dim3(256), 0, ctx->stream, X, n, val,
        pt.ptr, pt.col, pt.k, pt.e, Y, pt.ntile, nreal);
    prof_end(ctx, tk);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// Wide products (n >= kWideN): the tiles of k_tsgemm would re-read each
// column strip n/16 times from scattered 8-byte loads.  FP64: the library's
// dgemm (the reference's own DGEMM), split over K into batches (an n x n
// output alone is too few tiles for the chip when K is hundreds of
// thousands of rows) whose partial products are summed in FP64:
// strided-batched dgemm over chunks of kF64Chunk rows or more (at most
// kF64MaxParts partials) plus one dgemm for the remainder.  Split:
// k_tsgemm_split_wide (below).  A Gram (A == B) is mirrored from its lower
// triangle, so it is exactly symmetric.
constexpr int kWideN       = 256;
constexpr int kF64Chunk    = 16384;
constexpr int kF64MaxParts = 64;

// C = sum_c parts[c] (FP64, np partials of n x n); sym: both triangles from
// the lower
__global__ void k_sum_parts(const double* parts, int np, int n, int sym, double* C)
{
    const int64_t nn = (int64_t)n * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int col = (int)(t / n), row = (int)(t - (int64_t)col * n);
        if (sym && row < col) continue;
        double v = 0.;
        for (int c = 0; c < np; c++)
            v += parts[c * nn + t];
        C[t] = v;
        if (sym) C[(int64_t)row * n + col] = v;
    }
}

// The hi/lo split of a wide product on FP32 MFMA.  One block computes a
// kWT x kWT output tile (rows from A's columns i0.., columns from B's j0..)
// over one range of rows (a K-split).  Per chunk of kWK rows the block loads
// both column panels once, coalesced (a wave reads two columns' 512
// contiguous bytes per instruction), splits each value into hi = fl32(x),
// lo = fl32(x - hi) in registers and stores them to LDS column by column
// (stride kWL floats: the stores and the MFMA operand reads are free of bank
// conflicts).  Each wave owns a 32 x 32 quadrant (2 x 2 tiles of
// v_mfma_f32_16x16x4_f32): hi*hi and hi*lo + lo*hi accumulate in FP32 over
// the chunk's kWK rows, then the chunk sums are added in FP64.  Partial
// products of the K-splits go to part (FP64, n x n each, column-major).
constexpr int kWT = 64;  // output tile edge
constexpr int kWK = 64;  // rows per LDS panel
constexpr int kWR = 4;   // rows per FP32 sum of hi*hi (then added in FP64)
constexpr int kWL = 68;  // LDS column stride in floats (64 rows + 4)

__global__ void __launch_bounds__(256, 2) k_tsgemm_split_wide(const double* A, const double* B, int64_t ld,
    int64_t lnp, int n, int ntile, int64_t rows_per_split, int lower_only, double* part)
{
    typedef double dd2 __attribute__((ext_vector_type(2)));
    typedef float ff2 __attribute__((ext_vector_type(2)));
    // lower_only: the grid enumerates the tiles on and below the diagonal
    // only (so the round-robin of blocks over the XCDs stays balanced)
    const int tile = blockIdx.x;
    int ti, tj;
    if (lower_only)
    {
        ti = (int)((sqrtf(8.f * (float)tile + 1.f) - 1.f) * 0.5f);
        while ((ti + 1) * (ti + 2) / 2 <= tile)
            ti++;
        while (ti * (ti + 1) / 2 > tile)
            ti--;
        tj = tile - ti * (ti + 1) / 2;
    }
    else
    {
        ti = tile / ntile;
        tj = tile - ti * ntile;
    }
    const int split  = blockIdx.y;
    const int64_t k0 = (int64_t)split * rows_per_split;
    const int64_t k1 = min(k0 + rows_per_split, lnp);
    // a diagonal tile of a Gram reads one panel for both operands
    const bool one = A == B && ti == tj;
    __shared__ float Ah[kWT * kWL], Al[kWT * kWL], Bh[kWT * kWL], Bl[kWT * kWL];
    const float* Bhr = one ? Ah : Bh;
    const float* Blr = one ? Al : Bl;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wi = w & 1, wj = w >> 1;
    const int c = lane & 15, r = lane >> 4;
    const int i0 = ti * kWT, j0 = tj * kWT;
    // loader: rows 2 rr, 2 rr + 1 of column q * 8 + (t >> 5) for q = 0..7
    const int rr = 2 * (t & 31), cq = t >> 5;
    double acc[2][2][4] = {};
    for (int64_t kc = k0; kc < k1; kc += kWK)
    {
        const int64_t row = kc + rr;
#pragma unroll
        for (int q = 0; q < 8; q++)
        {
            const int col = q * 8 + cq;
            // two rows at once where they are 16-byte aligned (even ld),
            // else one by one
            auto ld2 = [&](const double* p) -> dd2 {
                dd2 v = { 0., 0. };
                if (row + 1 < k1 && (reinterpret_cast<uintptr_t>(p) & 15) == 0)
                    v = *reinterpret_cast<const dd2*>(p);
                else
                {
                    if (row < k1) v[0] = p[0];
                    if (row + 1 < k1) v[1] = p[1];
                }
                return v;
            };
            dd2 a = { 0., 0. }, b = { 0., 0. };
            if (i0 + col < n) a = ld2(A + (int64_t)(i0 + col) * ld + row);
            if (!one && j0 + col < n) b = ld2(B + (int64_t)(j0 + col) * ld + row);
            ff2 h, l;
            h[0] = (float)a[0];
            h[1] = (float)a[1];
            l[0] = (float)(a[0] - (double)h[0]);
            l[1] = (float)(a[1] - (double)h[1]);
            *reinterpret_cast<ff2*>(Ah + col * kWL + rr) = h;
            *reinterpret_cast<ff2*>(Al + col * kWL + rr) = l;
            if (!one)
            {
                h[0] = (float)b[0];
                h[1] = (float)b[1];
                l[0] = (float)(b[0] - (double)h[0]);
                l[1] = (float)(b[1] - (double)h[1]);
                *reinterpret_cast<ff2*>(Bh + col * kWL + rr) = h;
                *reinterpret_cast<ff2*>(Bl + col * kWL + rr) = l;
            }
        }
        __syncthreads();
        // hi*hi: FP32 sums over kWR rows (kWR / 4 MFMA steps), then into
        // FP64; the cross terms (2^-24 of it) in FP32 over the panel's rows
        f4 hx[2][2];
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++)
                hx[x][y] = f4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll 2
        for (int s0 = 0; s0 < kWK / 4; s0 += kWR / 4)
        {
            f4 hh[2][2];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
                    hh[x][y] = f4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int s = s0; s < s0 + kWR / 4; s++)
            {
                const int kk = 4 * s + r;
                float ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int x = 0; x < 2; x++)
                {
                    const int ia = (wi * 32 + x * 16 + c) * kWL + kk;
                    const int jb = (wj * 32 + x * 16 + c) * kWL + kk;
                    ah[x]        = Ah[ia];
                    al[x]        = Al[ia];
                    bh[x]        = Bhr[jb];
                    bl[x]        = Blr[jb];
                }
#pragma unroll
                for (int x = 0; x < 2; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++)
                    {
                        hh[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[x], bh[y], hh[x][y], 0, 0, 0);
                        hx[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[x], bl[y], hx[x][y], 0, 0, 0);
                        hx[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[x], bh[y], hx[x][y], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        acc[x][y][v] += (double)hh[x][y][v];
        }
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    acc[x][y][v] += (double)hx[x][y][v];
        __syncthreads();
    }
    // accumulator v of lane l: D[4 (l / 16) + v][l % 16] (FP32 MFMA layout)
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int row = i0 + wi * 32 + x * 16 + 4 * r + v;
                const int col = j0 + wj * 32 + y * 16 + c;
                if (row < n && col < n) part[((int64_t)split * n + col) * n + row] = acc[x][y][v];
            }
}

int tsgemm_wide(mgx_ctx* ctx, int prec, const double* A, const double* B, int64_t ld, int64_t lnp, int n,
    int subdivx, double vel, double* ss)
{
    const int64_t nn = (int64_t)n * n;
    const bool same  = A == B;
    rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
    for (int sl = 0; sl < subdivx; sl++)
    {
        const double* As = A + (int64_t)sl * lnp;
        const double* Bs = B + (int64_t)sl * lnp;
        double* C        = ss + sl * nn;
        if (prec == MGX_PREC_F64)
        {
            const double zero  = 0.;
            const int64_t kc   = std::max<int64_t>(kF64Chunk, (lnp + kF64MaxParts - 1) / kF64MaxParts);
            const int64_t nful = lnp / kc, rem = lnp - nful * kc;
            const int np       = (int)(nful + (rem > 0));
            double* parts      = (double*)workspace(ctx, 7, (size_t)np * nn * sizeof(double));
            if (!parts) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
            rocblas_status st = rocblas_status_success;
            // full products also for a Gram: rocBLAS dsyrk measured 10x slower here
            if (nful > 0)
                st = rocblas_dgemm_strided_batched(ctx->blas, rocblas_operation_transpose, rocblas_operation_none,
                    n, n, (rocblas_int)kc, &vel, As, (rocblas_int)ld, kc, Bs, (rocblas_int)ld, kc, &zero, parts, n,
                    nn, (rocblas_int)nful);
            if (st == rocblas_status_success && rem > 0)
                st = rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, n, n,
                    (rocblas_int)rem, &vel, As + nful * kc, (rocblas_int)ld, Bs + nful * kc, (rocblas_int)ld,
                    &zero, parts + nful * nn, n);
            if (st != rocblas_status_success) return set_err(ctx, MGX_EHIP, "rocBLAS FP64 product failed");
            hipLaunchKernelGGL(k_sum_parts, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, (const double*)parts, np,
                n, same ? 1 : 0, C);
            MGX_CHECK_LAUNCH(ctx);
            continue;
        }
        // split: K divided among enough blocks to fill the chip (at least
        // 32 chunks each), rows per split a multiple of the chunk
        const int ntile     = (n + kWT - 1) / kWT;
        const int active    = same ? ntile * (ntile + 1) / 2 : ntile * ntile;
        const int64_t nchk  = (lnp + kWK - 1) / kWK;
        const int64_t want  = std::max<int64_t>(1, (2048 + active - 1) / active);
        const int nsplit    = (int)std::max<int64_t>(1, std::min<int64_t>(want, nchk / 32));
        const int64_t rps   = ((nchk + nsplit - 1) / nsplit) * kWK;
        double* parts       = (double*)workspace(ctx, 7, (size_t)nsplit * nn * sizeof(double));
        if (!parts) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
        dim3 grid(active, nsplit), blk(256);
        hipLaunchKernelGGL(k_tsgemm_split_wide, grid, blk, 0, ctx->stream, As, Bs, ld, lnp, n, ntile, rps,
            same ? 1 : 0, parts);
        MGX_CHECK_LAUNCH(ctx);
        hipLaunchKernelGGL(k_chunk_reduce, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, (const double*)parts, n,
            nsplit, 1, vel, same ? 1 : 0, C);
        MGX_CHECK_LAUNCH(ctx);
    }
    return MGX_OK;
}

// part_ext: a caller-given buffer of part_cap bytes for the chunk partials
// (the number of chunks is lowered to fit it), else the context's workspace
int tsgemm(mgx_ctx* ctx, int prec, const double* A, const double* B, int64_t ld, int64_t lnp, int n, int subdivx,
    double vel, int lower_only, double* ss, double* part_ext = nullptr, size_t part_cap = 0)
{
    MGX_REQUIRE(ctx, A && B && ss && n > 0 && subdivx > 0 && lnp > 0, "bad product arguments");
    MGX_REQUIRE(ctx, (int64_t)subdivx * lnp <= ld, "slabs exceed the leading dimension");
    MGX_REQUIRE(ctx, prec == MGX_PREC_F64 || prec == MGX_PREC_SPLIT, "bad precision mode");
    if (n >= kWideN) return tsgemm_wide(ctx, prec, A, B, ld, lnp, n, subdivx, vel, ss);
    const int ntile  = (n + 15) / 16;
    const int ntiles = ntile * ntile;
    // enough chunks to fill the chip, each at least 256 rows
    int64_t want = (4096 + ntiles - 1) / ntiles;
    int64_t maxc = (lnp + 255) / 256;
    int nchunk   = (int)std::max<int64_t>(1, std::min<int64_t>(want, maxc));
    const size_t per_chunk = (size_t)subdivx * n * n * sizeof(double);
    double* part           = part_ext;
    if (part_ext)
    {
        MGX_REQUIRE(ctx, per_chunk <= part_cap, "subdivx * n * n * 8 bytes exceed the scratch buffer");
        nchunk = (int)std::min<size_t>((size_t)nchunk, part_cap / per_chunk);
    }
    const int64_t rows = (lnp + nchunk - 1) / nchunk;
    if (!part_ext)
    {
        part = (double*)workspace(ctx, 7, per_chunk * nchunk);
        if (!part) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    }
    dim3 grid(ntiles, nchunk, subdivx), blk(64 * kWaves);
    if (prec == MGX_PREC_F64)
        hipLaunchKernelGGL(k_tsgemm<MGX_PREC_F64>, grid, blk, 0, ctx->stream, A, B, ld, lnp, n, ntile, nchunk, rows,
            lower_only, part);
    else
        hipLaunchKernelGGL(k_tsgemm<MGX_PREC_SPLIT>, grid, blk, 0, ctx->stream, A, B, ld, lnp, n, ntile, nchunk,
            rows, lower_only, part);
    MGX_CHECK_LAUNCH(ctx);
    const int64_t tot = (int64_t)n * n * subdivx;
    hipLaunchKernelGGL(k_chunk_reduce, dim3(nblocks(tot)), dim3(kBlock), 0, ctx->stream, (const double*)part, n,
        nchunk, subdivx, vel, lower_only, ss);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int chol_inv(mgx_ctx* ctx, int n, const double* S, double* L, double* Sinv)
{
    MGX_REQUIRE(ctx, n > 0 && S && L && Sinv, "bad chol_inv arguments");
    int* info = (int*)workspace(ctx, 6, 64);
    if (!info) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    const size_t bytes = (size_t)n * n * sizeof(double);
    if (L != S) MGX_CHECK_HIP(ctx, hipMemcpyAsync(L, S, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (rocsolver_dpotrf(ctx->blas, rocblas_fill_lower, n, L, n, info) != rocblas_status_success)
        return set_err(ctx, MGX_EHIP, "rocsolver_dpotrf failed");
    int hinfo = 0;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hinfo != 0)
        return set_err(ctx, MGX_ENOTSPD, "ERROR in GramMatrix<MatrixType>::updateLS(): potrf info != 0");
    hipLaunchKernelGGL(k_tri, dim3(nblocks((int64_t)n * n)), dim3(kBlock), 0, ctx->stream, L, n, 1);
    MGX_CHECK_LAUNCH(ctx);
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(Sinv, L, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (rocsolver_dpotri(ctx->blas, rocblas_fill_lower, n, Sinv, n, info) != rocblas_status_success)
        return set_err(ctx, MGX_EHIP, "rocsolver_dpotri failed");
    hipLaunchKernelGGL(k_tri, dim3(nblocks((int64_t)n * n)), dim3(kBlock), 0, ctx->stream, Sinv, n, 0);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int diag_sum(mgx_ctx* ctx, int n, const double* A, double* dev_out)
{
    hipLaunchKernelGGL(k_diag_sum, dim3(1), dim3(256), 0, ctx->stream, A, n, dev_out);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// Refresh X ~ S^-1 by Newton-Schulz steps X <- X (2I - S X) from the previous
// inverse (S changes little between orbital steps, so the iteration starts
// well inside its quadratic basin).  Falls back to the Cholesky inverse when
// max|I - S X| does not start below 0.25 (first step, large change) or fails
// to contract.  work: 2 n^2 doubles.  *steps = Newton-Schulz steps taken
// (-1 when the Cholesky path ran).
int inverse_update(mgx_ctx* ctx, int n, const double* S, double* L, double* X, double* work, int* steps)
{
    const double one = 1., zero = 0., mone = -1., two = 2.;
    const size_t nn = (size_t)n * n;
    double* T  = work;
    double* X2 = work + nn;
    double* bm = (double*)workspace(ctx, 4, 1024 * sizeof(double));
    if (!bm) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
    double prev = 1e300;
    *steps      = 0;
    for (int it = 0; it < 8; it++)
    {
        if (rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, S, n, X, n, &zero,
                T, n)
            != rocblas_status_success)
            return set_err(ctx, MGX_EHIP, "rocblas_dgemm failed");
        const int nb = 512;
        hipLaunchKernelGGL(k_id_err, dim3(nb), dim3(256), 0, ctx->stream, (const double*)T, n, bm);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, ctx->stream, (const double*)bm, nb, bm + nb);
        MGX_CHECK_LAUNCH(ctx);
        double delta = 0.;
        MGX_CHECK_HIP(ctx, hipMemcpyAsync(&delta, bm + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (!(delta < 0.25) || delta > 0.5 * prev)
        {
            *steps = -1;
            return chol_inv(ctx, n, S, L, X);
        }
        if (delta < 1e-15) break;
        // X2 = 2 X - X T
        MGX_CHECK_HIP(ctx, hipMemcpyAsync(X2, X, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        if (rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &mone, X, n, T, n, &two,
                X2, n)
            != rocblas_status_success)
            return set_err(ctx, MGX_EHIP, "rocblas_dgemm failed");
        MGX_CHECK_HIP(ctx, hipMemcpyAsync(X, X2, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        (*steps)++;
        prev = delta;
        if (delta < 3e-8) break;  // the error after this step is ~delta^2
    }
    hipLaunchKernelGGL(k_symmetrize, dim3(nblocks((int64_t)nn)), dim3(kBlock), 0, ctx->stream, X, n);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// inverse_update for a pair-sparse S (values Sv per pair, column tiles pt):
// T = X S on the sparse product, X <- 2X - T X as one dense GEMM (written to
// the other half of the work buffer, operands swapped, no copies).  Stops
// before a step once delta = max|I - X S| < tol.  With FP64 products it also
// stops after a step once delta^2 (the error that step leaves) < tol; with
// lowp the correction E X, E = I - X S, is one FP32 product, whose rounding
// (~1e-7 |E| |X|) can exceed delta^2, so only measured deltas end the loop.
int inverse_update_pairs(mgx_ctx* ctx, int n, const double* S, const double* Sv, const PairTiles& pt, double* L,
    double* X, double* work, int* steps, double tol, bool lowp)
{
    const double mone = -1., two = 2.;
    const size_t nn = (size_t)n * n;
    double* T   = work;
    double* cur = X;
    double* nxt = work + nn;
    double* bm  = (double*)workspace(ctx, 4, 1024 * sizeof(double));
    if (!bm) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
    float *E32 = nullptr, *X32 = nullptr, *Y32 = nullptr;
    if (lowp)
    {
        E32 = (float*)workspace(ctx, 7, 3 * nn * sizeof(float));
        if (!E32) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
        X32 = E32 + nn;
        Y32 = X32 + nn;
    }
    const double gflop = 2. * (double)n * n * n;
    double prev = 1e300;
    *steps      = 0;
    int rc;
    for (int it = 0; it < 10; it++)
    {
        if ((rc = spmm_pairs(ctx, n, cur, Sv, pt, T))) return rc;
        const int nb = 512;
        if (lowp)
            hipLaunchKernelGGL(k_id_err_f32, dim3(nb), dim3(256), 0, ctx->stream, (const double*)T,
                (const double*)cur, n, bm, E32, X32);
        else
            hipLaunchKernelGGL(k_id_err, dim3(nb), dim3(256), 0, ctx->stream, (const double*)T, n, bm);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, ctx->stream, (const double*)bm, nb, bm + nb);
        MGX_CHECK_LAUNCH(ctx);
        double delta = 0.;
        MGX_CHECK_HIP(ctx, hipMemcpyAsync(&delta, bm + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (!(delta < 0.25) || delta > 0.5 * prev)
        {
            *steps = -1;
            return chol_inv(ctx, n, S, L, X);
        }
        if (delta < tol) break;
        if (lowp)
        {
            // nxt = cur + E cur
            const float onef = 1.f, zerof = 0.f;
            if (rocblas_sgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &onef, E32, n, X32,
                    n, &zerof, Y32, n)
                != rocblas_status_success)
                return set_err(ctx, MGX_EHIP, "rocblas_sgemm failed");
            hipLaunchKernelGGL(k_add_f32, dim3(nblocks((int64_t)nn)), dim3(kBlock), 0, ctx->stream,
                (const double*)cur, (const float*)Y32, nxt, (int64_t)nn);
            MGX_CHECK_LAUNCH(ctx);
        }
        else
        {
            // nxt = 2 cur - T cur
            const int tk = prof_begin(ctx, MGX_PROF_NSGEMM, gflop);
            if (rocblas_gemm_ex(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &mone, T,
                    rocblas_datatype_f64_r, n, cur, rocblas_datatype_f64_r, n, &two, cur, rocblas_datatype_f64_r, n,
                    nxt, rocblas_datatype_f64_r, n, rocblas_datatype_f64_r, rocblas_gemm_algo_standard, 0, 0)
                != rocblas_status_success)
                return set_err(ctx, MGX_EHIP, "rocblas_gemm_ex failed");
            prof_end(ctx, tk);
        }
        std::swap(cur, nxt);
        (*steps)++;
        prev = delta;
        if (!lowp && delta * delta < tol) break;  // the error after this step is ~delta^2
    }
    if (cur != X) MGX_CHECK_HIP(ctx, hipMemcpyAsync(X, cur, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_symmetrize, dim3(nblocks((int64_t)nn)), dim3(kBlock), 0, ctx->stream, X, n);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// inverse_update_pairs without host waits: nsteps Newton-Schulz steps from
// X, each recording delta_k = max|I - X_k S| (before step k) in deltas[k]
// on the device.  With bf16 the correction E X of a step is one BF16 product
// with K = 3n of hi/lo parts (Eh Xh + Eh Xl + El Xh, FP32 sums; relative
// error ~1e-5 of |E| |X|), so delta^2 does not bound the error a step leaves:
// the delta of the final X is measured too, into deltas[nsteps].  The caller
// reads the deltas later with its other scalars and checks them
// (ns_spec_check).
// the start of a refresh: X <- 2 X - P (on) and P <- the X it was given: the
// linear extrapolation of the last two inverses, S changing little from one
// step to the next (the steps then start an order closer)
__global__ void k_extrap(double* X, double* P, int64_t n, int on)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    {
        const double t = X[i];
        if (on) X[i] = 2. * t - P[i];
        P[i] = t;
    }
}

// xprev (may be null): the inverse of the step before, n x n; with extrap the
// steps start from 2 X - xprev, and xprev receives the X they were given
int inverse_update_spec(mgx_ctx* ctx, int n, const double* Sv, const PairTiles& pt, double* X, double* work,
    int nsteps, bool bf16, double* deltas, double* xprev, bool extrap)
{
    if (xprev)
    {
        hipLaunchKernelGGL(k_extrap, dim3(nblocks((int64_t)n * n)), dim3(kBlock), 0, ctx->stream, X, xprev,
            (int64_t)n * n, extrap ? 1 : 0);
        MGX_CHECK_LAUNCH(ctx);
    }
    const double mone = -1., two = 2.;
    const size_t nn = (size_t)n * n;
    double* T   = work;
    double* cur = X;
    double* nxt = work + nn;
    double* bm  = (double*)workspace(ctx, 4, 1024 * sizeof(double));
    if (!bm) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
    float* Y32     = nullptr;
    uint16_t* H16  = nullptr;
    if (bf16)
    {
        Y32 = (float*)workspace(ctx, 7, nn * sizeof(float) + 6 * nn * sizeof(uint16_t));
        if (!Y32) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
        H16 = reinterpret_cast<uint16_t*>(Y32 + nn);
    }
    const double flop = 2. * (double)n * n * n;
    int rc;
    const int nb = 512;
    for (int it = 0; it < nsteps; it++)
    {
        if ((rc = spmm_pairs(ctx, n, cur, Sv, pt, T))) return rc;
        if (bf16)
            hipLaunchKernelGGL(k_id_err_bf16, dim3(nb), dim3(256), 0, ctx->stream, (const double*)T,
                (const double*)cur, n, bm, H16, H16 + 3 * nn);
        else
            hipLaunchKernelGGL(k_id_err, dim3(nb), dim3(256), 0, ctx->stream, (const double*)T, n, bm);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, ctx->stream, (const double*)bm, nb, deltas + it);
        MGX_CHECK_LAUNCH(ctx);
        if (bf16)
        {
            const float onef = 1.f, zerof = 0.f;
            const int tk = prof_begin(ctx, MGX_PROF_NSGEMM16, 3. * flop);
            if (rocblas_gemm_ex(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, 3 * n, &onef, H16,
                    rocblas_datatype_bf16_r, n, H16 + 3 * nn, rocblas_datatype_bf16_r, 3 * n, &zerof, Y32,
                    rocblas_datatype_f32_r, n, Y32, rocblas_datatype_f32_r, n, rocblas_datatype_f32_r,
                    rocblas_gemm_algo_standard, 0, 0)
                != rocblas_status_success)
                return set_err(ctx, MGX_EHIP, "rocblas_gemm_ex (bf16) failed");
            prof_end(ctx, tk);
            hipLaunchKernelGGL(k_add_f32, dim3(nblocks((int64_t)nn)), dim3(kBlock), 0, ctx->stream,
                (const double*)cur, (const float*)Y32, nxt, (int64_t)nn);
            MGX_CHECK_LAUNCH(ctx);
        }
        else
        {
            const int tk = prof_begin(ctx, MGX_PROF_NSGEMM, flop);
            if (rocblas_gemm_ex(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &mone, T,
                    rocblas_datatype_f64_r, n, cur, rocblas_datatype_f64_r, n, &two, cur, rocblas_datatype_f64_r, n,
                    nxt, rocblas_datatype_f64_r, n, rocblas_datatype_f64_r, rocblas_gemm_algo_standard, 0, 0)
                != rocblas_status_success)
                return set_err(ctx, MGX_EHIP, "rocblas_gemm_ex failed");
            prof_end(ctx, tk);
        }
        std::swap(cur, nxt);
    }
    if (cur != X) MGX_CHECK_HIP(ctx, hipMemcpyAsync(X, cur, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_symmetrize_tiles, dim3((n + 31) / 32, (n + 31) / 32), dim3(256), 0, ctx->stream, X, n);
    MGX_CHECK_LAUNCH(ctx);
    if (bf16)
    {
        // the error the last BF16 step left, on the symmetrised X
        if ((rc = spmm_pairs(ctx, n, X, Sv, pt, T))) return rc;
        hipLaunchKernelGGL(k_id_err, dim3(nb), dim3(256), 0, ctx->stream, (const double*)T, n, bm);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, ctx->stream, (const double*)bm, nb, deltas + nsteps);
        MGX_CHECK_LAUNCH(ctx);
    }
    return MGX_OK;
}

// The checks inverse_update_pairs makes step by step, on the deltas of
// inverse_update_spec: 0 = the steps were enough (*needed = the steps the
// checked loop would have taken), 1 = they converge but more are needed,
// -1 = they do not converge (Cholesky).  With final (the BF16 steps)
// deltas[nsteps] holds the measured error of the final X, and only a
// measured delta below tol accepts.
int ns_spec_check(const double* deltas, int nsteps, double tol, bool final, int* needed)
{
    double prev = 1e300;
    for (int k = 0; k < nsteps; k++)
    {
        const double d = deltas[k];
        if (!(d < 0.25) || d > 0.5 * prev) return -1;
        if (d < tol)
        {
            *needed = k;
            return 0;
        }
        if (!final && d * d < tol)
        {
            *needed = k + 1;
            return 0;
        }
        prev = d;
    }
    if (final)
    {
        const double d = deltas[nsteps];
        if (!(d < 0.25) || d > 0.5 * prev) return -1;
        if (d < tol)
        {
            *needed = nsteps;
            return 0;
        }
    }
    *needed = nsteps;
    return 1;
}

int theta(mgx_ctx* ctx, int n, const double* Sinv, const double* H, double* th)
{
    const double one = 1., zero = 0.;
    rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
    if (rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, Sinv, n, H, n, &zero,
            th, n)
        != rocblas_status_success)
        return set_err(ctx, MGX_EHIP, "rocblas_dgemm failed");
    return MGX_OK;
}

int trace_prod(mgx_ctx* ctx, int n, const double* H, const double* Sinv, double* dev_out)
{
    hipLaunchKernelGGL(k_trace_prod, dim3(1), dim3(1024), 0, ctx->stream, H, Sinv, n, dev_out);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int ritz(mgx_ctx* ctx, int n, const double* H, const double* S, double* eig, double* Z)
{
    const size_t bytes = (size_t)n * n * sizeof(double);
    double* w          = (double*)workspace(ctx, 5, 2 * bytes + (size_t)n * sizeof(double) + 256);
    if (!w) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    double* A  = w;
    double* B  = w + (size_t)n * n;
    double* E  = B + (size_t)n * n;
    int* info  = (int*)workspace(ctx, 6, 64);
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(A, H, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(B, S, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (rocsolver_dsygvd(ctx->blas, rocblas_eform_ax, Z ? rocblas_evect_original : rocblas_evect_none,
            rocblas_fill_lower, n, A, n, B, n, eig, E, info)
        != rocblas_status_success)
        return set_err(ctx, MGX_EHIP, "rocsolver_dsygvd failed");
    int hinfo = 0;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hinfo != 0) return set_err(ctx, hinfo > n ? MGX_ENOTSPD : MGX_ENUMERIC, "sygvd did not converge / S not SPD");
    if (Z) MGX_CHECK_HIP(ctx, hipMemcpyAsync(Z, A, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return MGX_OK;
}

}  // namespace mgx

// Scratch of mgx_local_gram_noalloc: the chunk partials of k_tsgemm, at most
// kNoAllocBytes.  The array is the device memory itself (one copy per GPU,
// held until the process ends); host code uses only the device address that
// scratch_init() takes from hipGetSymbolAddress, never the array's name.
constexpr size_t kNoAllocBytes = size_t(16) << 20;
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

namespace mgx
{
namespace
{
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// once per device, with that device current (context creation): the device
// address of this GPU's copy of `scratch`
int scratch_init(mgx_ctx* ctx, int device)
{
    if (device < 0 || device >= kMaxDevices) return set_err(ctx, MGX_EINVAL, "device index out of range for scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return MGX_OK;
    void* p        = nullptr;
    hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p)
        return set_err(ctx, MGX_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(err));
    g_scratch_ptr[device] = p;
    return MGX_OK;
}

}  // namespace mgx

using namespace mgx;

extern "C" {

int mgx_local_gram(mgx_ctx* ctx, int prec, const double* psi, int64_t ld, int64_t lnp, int n, int subdivx,
    double vel, double* ss)
{
    int rc = tsgemm(ctx, prec, psi, psi, ld, lnp, n, subdivx, vel, 1, ss);
    return rc ? rc : sync(ctx);
}

int mgx_local_gram_noalloc(mgx_ctx* ctx, int prec, const double* psi, int64_t ld, int64_t lnp, int n, int subdivx,
    double vel, double* ss)
{
    if (!ctx) return MGX_EINVAL;
    MGX_REQUIRE(ctx, n > 0 && n < kWideN, "mgx_local_gram_noalloc takes n < 256 (wider products use rocBLAS)");
    void* p = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_ptr[ctx->device] : nullptr;
    MGX_REQUIRE(ctx, p, "scratch address of this device not set up (mgx_ctx_create)");
    // every partial the reduction reads is written by the same call: nothing
    // to clear; calls that overlap on one GPU share the array and race
    int rc = tsgemm(ctx, prec, psi, psi, ld, lnp, n, subdivx, vel, 1, ss, (double*)p, kNoAllocBytes);
    return rc ? rc : sync(ctx);
}

int mgx_local_product(mgx_ctx* ctx, int prec, const double* a, const double* b, int64_t ld, int64_t lnp, int n,
    int subdivx, double vel, double* ss)
{
    int rc = tsgemm(ctx, prec, a, b, ld, lnp, n, subdivx, vel, 0, ss);
    return rc ? rc : sync(ctx);
}

int mgx_local_to_global(mgx_ctx* ctx, const double* ss, const int* gids, int subdivx, int m, int numst, double tol,
    double* glob)
{
    MGX_REQUIRE(ctx, ss && gids && glob && subdivx > 0 && m > 0 && numst > 0, "bad local_to_global arguments");
    hipLaunchKernelGGL(k_l2g, dim3(nblocks((int64_t)m * m)), dim3(kBlock), 0, ctx->stream, ss, gids, subdivx, m,
        numst, tol, glob);
    MGX_CHECK_LAUNCH(ctx);
    return sync(ctx);
}

int mgx_chol_inv(mgx_ctx* ctx, int n, const double* S, double* L, double* Sinv)
{
    int rc = chol_inv(ctx, n, S, L, Sinv);
    return rc ? rc : sync(ctx);
}

int mgx_theta(mgx_ctx* ctx, int n, const double* Sinv, const double* H, double* th)
{
    MGX_REQUIRE(ctx, n > 0 && Sinv && H && th, "bad theta arguments");
    int rc = theta(ctx, n, Sinv, H, th);
    return rc ? rc : sync(ctx);
}

int mgx_eigsum(mgx_ctx* ctx, int n, const double* H, const double* Sinv, double* trace)
{
    MGX_REQUIRE(ctx, n > 0 && H && Sinv && trace, "bad eigsum arguments");
    double* d = (double*)workspace(ctx, 4, 256);
    if (!d) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    int rc = trace_prod(ctx, n, H, Sinv, d);
    if (rc) return rc;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(trace, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return sync(ctx);
}

int mgx_ritz(mgx_ctx* ctx, int n, const double* H, const double* S, double* eig, double* Z)
{
    MGX_REQUIRE(ctx, n > 0 && H && S && eig, "bad ritz arguments");
    int rc = ritz(ctx, n, H, S, eig, Z);
    return rc ? rc : sync(ctx);
}

int mgx_mehr_projected(mgx_ctx* ctx, int n, const double* S, const double* Sinv, const double* B, const double* H,
    double* invB, double* theta, double* HB, double* eigsum)
{
    MGX_REQUIRE(ctx, n > 0 && S && Sinv && B && H && invB && theta && HB, "bad projected-matrix arguments");
    const size_t nn = (size_t)n * n;
    double* w       = (double*)workspace(ctx, 5, 2 * nn * sizeof(double) + 64);
    if (!w) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    // invB (ProjectedMatricesMehrstellen::computeInvB: potrf + potri)
    int rc = chol_inv(ctx, n, B, w, invB);
    if (rc) return rc;
    // theta = invB H (updateTheta)
    if ((rc = mgx::theta(ctx, n, invB, H, theta))) return rc;
    // HB = 0.5 (S theta + (S theta)^T) (updateHB)
    const double one = 1., zero = 0.;
    rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
    if (rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, S, n, theta, n, &zero,
            w, n)
        != rocblas_status_success)
        return set_err(ctx, MGX_EHIP, "rocblas_dgemm failed");
    hipLaunchKernelGGL(k_sym_half, dim3(nblocks((int64_t)nn)), dim3(kBlock), 0, ctx->stream, (const double*)w, n, HB);
    MGX_CHECK_LAUNCH(ctx);
    // eigsum = tr(HB Sinv) (getEigSum)
    if (eigsum)
    {
        double* dv = w + nn;
        if ((rc = trace_prod(ctx, n, HB, Sinv, dv))) return rc;
        MGX_CHECK_HIP(ctx, hipMemcpyAsync(eigsum, dv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    return sync(ctx);
}

int mgx_residual(mgx_ctx* ctx, const double* psi, const double* hpsi, const double* th, int64_t ld, int64_t numpt,
    int n, double* res)
{
    MGX_REQUIRE(ctx, psi && hpsi && th && res && n > 0 && numpt > 0 && ld >= numpt, "bad residual arguments");
    const double one = 1., zero = 0.;
    rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
    if (rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)numpt, n, n, &one, psi,
            (rocblas_int)ld, th, n, &zero, res, (rocblas_int)ld)
        != rocblas_status_success)
        return set_err(ctx, MGX_EHIP, "rocblas_dgemm failed");
    hipLaunchKernelGGL(k_sub, dim3(nblocks(numpt * n)), dim3(kBlock), 0, ctx->stream, res, hpsi, ld, numpt, n);
    MGX_CHECK_LAUNCH(ctx);
    return sync(ctx);
}

int mgx_residual_norm(mgx_ctx* ctx, const double* res, int64_t ld, int64_t numpt, int n, double vel,
    const double* Sinv, double* norm)
{
    MGX_REQUIRE(ctx, res && Sinv && norm && n > 0, "bad residual_norm arguments");
    double* d = (double*)workspace(ctx, 4, (n + 8) * sizeof(double));
    if (!d) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    hipLaunchKernelGGL(k_col_sumsq, dim3(n), dim3(256), 0, ctx->stream, res, ld, numpt, vel, d);
    hipLaunchKernelGGL(k_diag_dot, dim3(1), dim3(256), 0, ctx->stream, (const double*)d, Sinv, n, d + n);
    MGX_CHECK_LAUNCH(ctx);
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(norm, d + n, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return sync(ctx);
}

}  // extern "C"
