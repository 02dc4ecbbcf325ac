// box_ops.hip -- per-orbital cubic boxes on a periodic grid.
//
// A box is the reference's localisation mask (Map2Masks / GridMaskMult,
// src/Map2Masks.cc:26-58) specialised to a 0/1 cube: data outside the box is
// zero, so stencils read zeros there and every result is stored on the box
// only (the mask applied after each operation).  Storage: box j at level l is
// (edge_j >> l)^3 values, x slowest, packed one box after another.
//
// Work decomposition: tiles of 256 consecutive points of one box (one point
// per lane), blocks remapped so that consecutive tiles of a box run on the
// same XCD and share its L2.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

struct mgx_boxset
{
    mgx_ctx* ctx = nullptr;
    int nbox = 0, levels = 0;
    int gdim[3];
    double h[3];
    std::vector<int> origin, edge;
    std::vector<std::vector<int64_t>> off;  // [level][nbox+1]
    int* d_origin = nullptr;
    int* d_edge   = nullptr;
    std::vector<int64_t*> d_off;
    std::vector<int> ntiles;
    std::vector<int*> d_tile;  // per level: 2*ntiles ints (box, start)
    // interacting pairs i <= j (overlap or within the stencil reach): 11 ints
    // (i, j, a[3], b[3], sh[3]); [a,b) = overlap in i-local (may be empty),
    // sh maps i-local -> j-local
    int64_t npairs = 0;
    int* d_pairs   = nullptr;
    // neighbours of each box (k ascending, incl. itself): 10 ints
    // (k, a[3], b[3], sh[3]) in j-local coordinates, sh maps j-local -> k-local
    int* d_nbr_ptr = nullptr;
    int* d_nbr     = nullptr;
    int64_t nnbr   = 0;
    int maxedge    = 0;
    // spatial cells of 8^3 points (level 0) listing the boxes that cover
    // them: 10 ints per entry (box, cell origin in box-local coords [3],
    // covered cell-local range lo[3], hi[3]); boxes ascending in each cell
    int ncell = 0, kmax = 0;
    int* d_cell_ptr  = nullptr;
    int* d_cell_ent  = nullptr;
    void* d_cell_ce  = nullptr;  // the entries as the kernels read them (mgx::CellEnt)
    int64_t* d_cell_slot = nullptr;  // base of the cell's k*k partial block
    int64_t nslot    = 0;
    // per pair: the partial slots (ab, ba) of every cell both boxes cover
    int* d_pslot_ptr = nullptr;
    int64_t* d_pslot = nullptr;
    // per box: its cell entries in cell order (sums of per-entry partials)
    int* d_bent_ptr  = nullptr;
    int* d_bent      = nullptr;
    int nent         = 0;  // cell entries in all
    bool use_cells   = false;
    bool pair_z      = false;  // every origin and edge even: paired z loads in the cell kernels
    // per box j: the other boxes i it interacts with (overlap or within 2
    // points), 4 ints (i, sh[3]) with sh mapping j-local -> i-local; and per
    // pair the entries (i in j's list, j in i's list) for the skin terms
    int* d_ilist_ptr = nullptr;
    int* d_ilist     = nullptr;
    int64_t nilist   = 0;
    int* d_pair_ient = nullptr;  // 2 ints per pair (-1 for i == j)
    // the skin's algorithmic bytes per call (k_slab_skin): for every face of
    // box j with a box i < j across it, the two planes of psi_j next to the
    // face, and per such i the psi_i values on the two layers beyond the
    // face inside box i
    int64_t skin_bytes = 0;
    // column tiles of the pair-sparse S / H for sparse x dense products
    int ntile        = 0;
    int* d_sp_ptr    = nullptr;
    int* d_sp_col    = nullptr;
    int* d_sp_k      = nullptr;
    int* d_sp_e      = nullptr;
    // autotuned launch shapes (MGX_TUNE_*) and whether the tuner ran or a
    // caller set them
    int tune[MGX_TUNE_N] = { 1, 1, 1, 1, 1, 1, 1, 1 };
    bool tuned           = false;
};

namespace mgx
{

constexpr int kTile = 256;

namespace
{

constexpr double kInv12 = 1. / 12.;
constexpr double kInv64 = 1. / 64.;

int wrap_h(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// unique periodic shift (i-local -> j-local) under which box i meets box j
// extended by the stencil reach (2 points); false if they do not interact
bool ext_shift1d(int oi, int ei, int oj, int ej, int G, int* sh)
{
    const int s = wrap_h(oj - oi, G);
    for (int img = -1; img <= 1; img++)
    {
        const int shift = -s + img * G;
        const int lo = shift, hi = ei - 1 + shift;
        if (hi >= -2 && lo <= ej + 1)
        {
            *sh = shift;
            return true;
        }
    }
    return false;
}

struct BoxView
{
    const int64_t* off;
    const int* edge;
    const int* origin;
    const int* tile;
    int level;
    int ntiles;
    int g0, g1, g2;  // global dims at this level
};

BoxView view(const mgx_boxset* bs, int level)
{
    BoxView v;
    v.off    = bs->d_off[level];
    v.edge   = bs->d_edge;
    v.origin = bs->d_origin;
    v.tile   = bs->d_tile[level];
    v.level  = level;
    v.ntiles = bs->ntiles[level];
    v.g0     = bs->gdim[0] >> level;
    v.g1     = bs->gdim[1] >> level;
    v.g2     = bs->gdim[2] >> level;
    return v;
}

unsigned tile_grid(int ntiles) { return (unsigned)(((ntiles + 7) / 8) * 8); }

// resolve the tile of this block; false if the block is padding
__device__ __forceinline__ bool tile_point(const BoxView& v, int* j, int* e, int* x, int* y, int* z, int64_t* base,
    int* p)
{
    const int64_t t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= v.ntiles) return false;
    *j          = v.tile[2 * t];
    const int s = v.tile[2 * t + 1];
    *e          = v.edge[*j] >> v.level;
    const int n = (*e) * (*e) * (*e);
    *p          = s + threadIdx.x;
    if (*p >= n) return false;
    *base  = v.off[*j];
    int r  = *p;
    *x     = r / ((*e) * (*e));
    r -= (*x) * (*e) * (*e);
    *y = r / (*e);
    *z = r - (*y) * (*e);
    return true;
}

template <typename T>
__device__ __forceinline__ T bv(const T* a, int e, int x, int y, int z)
{
    if ((unsigned)x >= (unsigned)e || (unsigned)y >= (unsigned)e || (unsigned)z >= (unsigned)e) return (T)0;
    return a[((int64_t)x * e + y) * e + z];
}

// stencils on a box (zero outside) -- same arithmetic as the ghosted kernels
template <typename T, int TYPE>
__device__ __forceinline__ double box_stencil(const T* a, int e, int x, int y, int z, const Coef4& c4,
    const Coef2& c2, const CoefM& cm, double diag4)
{
#define A(dx, dy, dz) bv(a, e, x + (dx), y + (dy), z + (dz))
    if (TYPE == MGX_LAP_4TH)
        return diag4 * (double)A(0, 0, 0) + c4.c1x * ((double)A(-1, 0, 0) + (double)A(1, 0, 0))
               + c4.c1y * ((double)A(0, -1, 0) + (double)A(0, 1, 0))
               + c4.c1z * ((double)A(0, 0, -1) + (double)A(0, 0, 1))
               + c4.c2x * ((double)A(-2, 0, 0) + (double)A(2, 0, 0))
               + c4.c2y * ((double)A(0, -2, 0) + (double)A(0, 2, 0))
               + c4.c2z * ((double)A(0, 0, -2) + (double)A(0, 0, 2));
    else if (TYPE == MGX_LAP_2ND)
        return c2.c0 * A(0, 0, 0) + c2.c1x * (T)(A(-1, 0, 0) + A(1, 0, 0)) + c2.c1y * (T)(A(0, -1, 0) + A(0, 1, 0))
               + c2.c1z * (T)(A(0, 0, -1) + A(0, 0, 1));
    else
        return cm.c0 * (double)A(0, 0, 0) + cm.cz * (double)(T)(A(0, 0, -1) + A(0, 0, 1))
               + cm.cy * (double)(T)(A(0, -1, 0) + A(0, 1, 0)) + cm.cx * (double)(T)(A(-1, 0, 0) + A(1, 0, 0))
               + cm.cxz * (double)(T)(A(-1, 0, -1) + A(-1, 0, 1) + A(1, 0, -1) + A(1, 0, 1))
               + cm.cyz * (double)(T)(A(0, -1, -1) + A(0, -1, 1) + A(0, 1, -1) + A(0, 1, 1))
               + cm.cxy * (double)(T)(A(-1, -1, 0) + A(1, -1, 0) + A(-1, 1, 0) + A(1, 1, 0));
#undef A
}

struct Coefs
{
    Coef4 c4;
    Coef2 c2;
    CoefM cm;
};

Coefs coefs_at(const mgx_boxset* bs, int level)
{
    double h[3] = { bs->h[0] * (1 << level), bs->h[1] * (1 << level), bs->h[2] * (1 << level) };
    Coefs c;
    c.c4 = coef4(h);
    c.c2 = coef2(h);
    c.cm = coefm(h);
    return c;
}

template <typename T, int TYPE>
__global__ void __launch_bounds__(256) k_box_lap(const T* in, T* out, BoxView v, Coefs c)
{
    int j, e, x, y, z, p;
    int64_t base;
    if (!tile_point(v, &j, &e, &x, &y, &z, &base, &p)) return;
    out[base + p] = (T)box_stencil<T, TYPE>(in + base, e, x, y, z, c.c4, c.c2, c.cm, c.c4.c0);
}

// FDoper::del2_4th_withPot on a box, V gathered from the periodic grid
template <typename T>
__global__ void __launch_bounds__(256) k_box_hloc(const T* in, const double* vtot, T* out, BoxView v, Coef4 c4)
{
    int j, e, x, y, z, p;
    int64_t base;
    if (!tile_point(v, &j, &e, &x, &y, &z, &base, &p)) return;
    int gx = v.origin[3 * j] + x, gy = v.origin[3 * j + 1] + y, gz = v.origin[3 * j + 2] + z;
    if (gx >= v.g0) gx -= v.g0;
    if (gy >= v.g1) gy -= v.g1;
    if (gz >= v.g2) gz -= v.g2;
    const double pot = vtot[((int64_t)gx * v.g1 + gy) * v.g2 + gz];
    Coef2 c2{};
    CoefM cm{};
    out[base + p] = (T)box_stencil<T, MGX_LAP_4TH>(in + base, e, x, y, z, c4, c2, cm, c4.c0 + pot);
}

// One GridFuncVector::jacobi sweep + mask, out of place:
// w = (T)L x; w -= f; x_out = x + (T)(-jf*w)
template <typename T, int TYPE>
__global__ void __launch_bounds__(256) k_box_jacobi(const T* x, const T* f, T* xo, T* w, BoxView v, Coefs c,
    double mjf)
{
    int j, e, xx, y, z, p;
    int64_t base;
    if (!tile_point(v, &j, &e, &xx, &y, &z, &base, &p)) return;
    const int64_t i = base + p;
    T wv            = (T)box_stencil<T, TYPE>(x + base, e, xx, y, z, c.c4, c.c2, c.cm, c.c4.c0);
    wv += static_cast<T>(-1. * static_cast<double>(f[i]));
    if (w) w[i] = wv;
    T xv = x[i];
    xv += static_cast<T>(mjf * static_cast<double>(wv));
    xo[i] = xv;
}

// ---------------------------------------------------------------------------
// Plane-marching stencil sweeps (2nd / 4th order) on boxes.  One workgroup per
// box walks x; each thread owns P points of the y-z plane and keeps their
// values at x-2..x+2 in a register queue, while the current plane sits in LDS
// with a zero ring of width 2 (the box boundary).  Each input element is read
// from HBM once; the next plane is prefetched before the current one is
// computed.  The arithmetic per point is the reference's, in the same order.
//   MODE 0: out = (T) L x
//   MODE 1: out = (T)((c0 + V) x + neighbours)   (del2_4th_withPot)
//   MODE 2: one Jacobi sweep: w = (T)L x; w -= f; xo = x + (T)(mjf*w)
enum
{
    MARCH_LAP    = 0,
    MARCH_HLOC   = 1,
    MARCH_JACOBI = 2
};

struct MarchArgs
{
    const int64_t* off;
    const int* edge;
    const int* origin;
    int level;
    int g0, g1, g2;
    int lds_w;  // LDS row width = max edge + 4
    double mjf;
    // HLOC only: a pending normalisation psi_j *= 1/sqrt(nrm[j]), applied as
    // psi is loaded (and written back to psi_w, the same array as in)
    const double* nrm = nullptr;
    double* psi_w     = nullptr;
    // JACOBI only, last fine sweep of an Anderson step: per-box sums of
    // (f - f0)^2 and (f - f0) f with f = da * x_out, f0 = da * dkr0 (type T),
    // times dvel, to dout[j] and dout[nbox + j] (AndersonMix::update's dots)
    const void* dkr0 = nullptr;
    double da = 0., dvel = 0.;
    double* dout = nullptr;
    // JACOBI with an output type other than T (the FP16 V-cycle's last fine
    // sweep): x_out = x * oinv[j], undoing the box's scale of the residual;
    // dkr0 then has the output type
    const double* oinv = nullptr;
    // the boxes in Morton order of their centres (nbox of them), blocks
    // mapped XCD-aware: nearby boxes, whose H psi reads overlapping parts of
    // V, run on one XCD
    const int* order = nullptr;
    int nbox         = 0;
    // JACOBI (k_box_march4, no dots, output type T): the stale residual w
    // restricted to the next level inside the sweep (MGkernelRestrict3D,
    // zero outside the box) and written to rc (type T) at the coarse level's
    // offsets offc, instead of w itself
    void* rc            = nullptr;
    const int64_t* offc = nullptr;
    // with rc: the coarse level's zero-start first sweep x = 0 + rjf f
    // (k_first_sweep's arithmetic) written to rx beside it
    void* rx   = nullptr;
    double rjf = 0.;
    // JACOBI (k_box_march4, no dots, output type T), first sweep after the
    // coarse correction: the input is in - P(pc) (MGkernelExtend3D of the
    // coarse correction pc, type T, at offsets offc), formed as each plane
    // is loaded, instead of being written back by a separate prolongation;
    // cpl = elements per coarse plane in LDS, (max coarse edge + 1)^2
    const void* pc = nullptr;
    int cpl        = 0;
};

// the box of this block (grids padded to a multiple of 8); -1: no box
__device__ __forceinline__ int march_box(const MarchArgs& a)
{
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    return L < a.nbox ? a.order[L] : -1;
}

struct MarchDots
{
    const void* kr0;
    double a, vel;
    double* out;
};

// NB: number of LDS plane buffers.  With 2 the planes alternate and one
// barrier per plane suffices (plane x+1 is written to the buffer that plane
// x-1 was read from, and every thread has passed plane x's barrier, so it is
// done reading plane x-1).  PD: planes loaded ahead (1: plane x+3 is loaded
// while plane x is computed; 2: plane x+4).
template <typename T, int TYPE, int MODE, int NT, int P, int NB, int PD>
__global__ void __launch_bounds__(NT) k_box_march(const T* __restrict__ in, const T* __restrict__ f,
    const double* __restrict__ vtot, T* __restrict__ out, T* __restrict__ wout, MarchArgs a, Coef4 c4, Coef2 c2)
{
    extern __shared__ unsigned char lds_raw[];
    T* pl0      = reinterpret_cast<T*>(lds_raw);
    const int j = march_box(a);
    if (j < 0) return;
    const int e = a.edge[j] >> a.level;
    const int E2 = e * e;
    const int W  = a.lds_w;
    const int64_t base = a.off[j];
    // zero the plane buffers including their rings once; the rings are never
    // written again
    for (int t = threadIdx.x; t < NB * W * W; t += NT)
        pl0[t] = (T)0;
    int qy[P], qz[P];
    bool act[P];
#pragma unroll
    for (int k = 0; k < P; k++)
    {
        const int q = threadIdx.x + k * NT;
        act[k]      = q < E2;
        qy[k]       = act[k] ? q / e : 0;
        qz[k]       = act[k] ? q - (q / e) * e : 0;
    }
    // the pending normalisation (HLOC on FP64 only): every value is loaded
    // once, scaled with LocGridOrbitals::normalize's factor and written back
    const bool scl  = MODE == MARCH_HLOC && sizeof(T) == sizeof(double) && a.nrm != nullptr;
    const double fs = scl ? 1. / sqrt(a.nrm[j]) : 1.;
    auto ld = [&](int64_t idx) -> T {
        if (!scl) return in[idx];
        double v = a.psi_w[idx];
        v *= fs;
        a.psi_w[idx] = v;
        return (T)v;
    };
    // qv[k][d]: plane x-2+d at this thread's point; nx[i]: plane x+3+i
    T qv[P][5];
    T nx[PD][P];
#pragma unroll
    for (int k = 0; k < P; k++)
    {
        const int q = threadIdx.x + k * NT;
        qv[k][0]    = (T)0;
        qv[k][1]    = (T)0;
        qv[k][2]    = act[k] ? ld(base + q) : (T)0;
        qv[k][3]    = (act[k] && e > 1) ? ld(base + E2 + q) : (T)0;
        qv[k][4]    = (act[k] && e > 2) ? ld(base + 2 * E2 + q) : (T)0;
#pragma unroll
        for (int i = 0; i + 1 < PD; i++)
            nx[i][k] = (act[k] && e > 3 + i) ? ld(base + (int64_t)(3 + i) * E2 + q) : (T)0;
    }
    __syncthreads();
    for (int x = 0; x < e; x++)
    {
        T* pl = (NB == 2 && (x & 1)) ? pl0 + W * W : pl0;
        // load plane x+2+PD; the potential of plane x (HLOC) is gathered here
        // too, so its latency overlaps the LDS store and the barrier
        double pot[P];
#pragma unroll
        for (int k = 0; k < P; k++)
        {
            const int q = threadIdx.x + k * NT;
            const int xn = x + 2 + PD;
            nx[PD - 1][k] = (act[k] && xn < e) ? ld(base + (int64_t)xn * E2 + q) : (T)0;
            pot[k]        = 0.;
            if (MODE == MARCH_HLOC && TYPE == MGX_LAP_4TH && act[k])
            {
                int gx = a.origin[3 * j] + x, gy = a.origin[3 * j + 1] + qy[k], gz = a.origin[3 * j + 2] + qz[k];
                if (gx >= a.g0) gx -= a.g0;
                if (gy >= a.g1) gy -= a.g1;
                if (gz >= a.g2) gz -= a.g2;
                pot[k] = vtot[((int64_t)gx * a.g1 + gy) * a.g2 + gz];
            }
        }
#pragma unroll
        for (int k = 0; k < P; k++)
            if (act[k]) pl[(qy[k] + 2) * W + qz[k] + 2] = qv[k][2];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < P; k++)
        {
            if (!act[k]) continue;
            const int q   = threadIdx.x + k * NT;
            const T* ctr  = pl + (qy[k] + 2) * W + qz[k] + 2;
            const int64_t gi = base + (int64_t)x * E2 + q;
            double r;
            if (TYPE == MGX_LAP_4TH)
            {
                double diag = c4.c0;
                if (MODE == MARCH_HLOC) diag = c4.c0 + pot[k];
                r = diag * (double)qv[k][2] + c4.c1x * ((double)qv[k][1] + (double)qv[k][3])
                    + c4.c1y * ((double)ctr[-W] + (double)ctr[W]) + c4.c1z * ((double)ctr[-1] + (double)ctr[1])
                    + c4.c2x * ((double)qv[k][0] + (double)qv[k][4])
                    + c4.c2y * ((double)ctr[-2 * W] + (double)ctr[2 * W])
                    + c4.c2z * ((double)ctr[-2] + (double)ctr[2]);
            }
            else
            {
                r = c2.c0 * qv[k][2] + c2.c1x * (T)(qv[k][1] + qv[k][3]) + c2.c1y * (T)(ctr[-W] + ctr[W])
                    + c2.c1z * (T)(ctr[-1] + ctr[1]);
            }
            if (MODE == MARCH_JACOBI)
            {
                T wv = (T)r;
                wv += static_cast<T>(-1. * static_cast<double>(f[gi]));
                if (wout) wout[gi] = wv;
                T xv = qv[k][2];
                xv += static_cast<T>(a.mjf * static_cast<double>(wv));
                out[gi] = xv;
            }
            else
                out[gi] = (T)r;
        }
        if (NB == 1) __syncthreads();
#pragma unroll
        for (int k = 0; k < P; k++)
        {
            qv[k][0] = qv[k][1];
            qv[k][1] = qv[k][2];
            qv[k][2] = qv[k][3];
            qv[k][3] = qv[k][4];
            qv[k][4] = nx[0][k];
#pragma unroll
            for (int i = 0; i + 1 < PD; i++)
                nx[i][k] = nx[i + 1][k];
        }
    }
}

template <typename T>
__device__ __forceinline__ T prolong8_val(const T (&c)[2][2][2], int ox, int oy, int oz);

// The same sweeps with V consecutive z-points per thread (V = 4 for FP32,
// 2 for FP64: 16 bytes; every box edge at this level a multiple of 4): one
// 16-byte load or store per array and plane instead of V, and the y- and
// z-neighbours read from LDS as whole V-wide pieces.  LDS rows have a ring of
// ZR = max(V, 2) in z (so the pieces stay aligned) and 2 in y; the arithmetic
// per point is that of k_box_march.
template <typename T, int TYPE, int MODE, int NT, int P, int NB, int V, int XT, typename TO = T>
__global__ void __launch_bounds__(NT, XT ? 6 : 8) k_box_march4(const T* __restrict__ in, const T* __restrict__ f,
    const double* __restrict__ vtot, TO* __restrict__ out, T* __restrict__ wout, MarchArgs a, Coef4 c4, Coef2 c2)
{
    typedef T V4 __attribute__((ext_vector_type(V)));
    typedef TO VO __attribute__((ext_vector_type(V)));
    constexpr bool mixed = !std::is_same<T, TO>::value;
    typedef T V2 __attribute__((ext_vector_type(2)));
    constexpr int ZR = V > 2 ? V : 2;
    extern __shared__ unsigned char lds_raw[];
    T* pl0      = reinterpret_cast<T*>(lds_raw);
    const int j = march_box(a);
    if (j < 0) return;
    const int e = a.edge[j] >> a.level;
    const int E2 = e * e;
    const int W  = a.lds_w;  // max edge + 2 ZR
    const int HW = W * (W - 2 * ZR + 4);
    const int64_t base = a.off[j];
    // XT 3: three coarse planes after the stencil planes, zero in their last
    // row and column (coarse index ec, outside the box)
    constexpr bool pro = XT == 3 && MODE == MARCH_JACOBI && std::is_same<T, TO>::value;
    for (int t = threadIdx.x; t < NB * HW + (pro ? 3 * a.cpl : 0); t += NT)
        pl0[t] = (T)0;
    int qy[P], qz[P];
    bool act[P];
#pragma unroll
    for (int k = 0; k < P; k++)
    {
        const int q = V * (threadIdx.x + k * NT);
        act[k]      = q < E2;
        qy[k]       = act[k] ? q / e : 0;
        qz[k]       = act[k] ? q - (q / e) * e : 0;
    }
    const bool scl  = MODE == MARCH_HLOC && sizeof(T) == sizeof(double) && a.nrm != nullptr;
    const double fs = scl ? 1. / sqrt(a.nrm[j]) : 1.;
    auto ld = [&](int64_t idx) -> V4 {
        if (!scl) return *reinterpret_cast<const V4*>(in + idx);
        V4 v = *reinterpret_cast<const V4*>(a.psi_w + idx);
#pragma unroll
        for (int i = 0; i < V; i++)
        {
            double d = (double)v[i];
            d *= fs;
            v[i] = (T)d;
        }
        *reinterpret_cast<V4*>(a.psi_w + idx) = v;
        return v;
    };
    V4 zero4;
#pragma unroll
    for (int i = 0; i < V; i++)
        zero4[i] = (T)0;
    V4 qv[P][5];
    V4 nx[P];
    // XT: 0 the sweep alone, 1 with the Anderson dots, 2 with the restriction
    constexpr bool dots = XT == 1 && MODE == MARCH_JACOBI;
    const TO* dkr0  = reinterpret_cast<const TO*>(a.dkr0);
    const double oj = mixed ? a.oinv[j] : 1.;
    double dacc0 = 0., dacc1 = 0.;
    // the fused restriction: each plane of w goes to a ring of two e x e
    // planes in LDS behind the stencil planes; after the next barrier every
    // thread forms the 3 x 3 (y, z) sums with weights (1 2 1) x (1 2 1) of
    // its coarse points (ec^2 <= 2 NT: the host gives NT P V >= e^2 and
    // V P <= 8) and adds them to the coarse plane in registers with the x
    // weights 1 2 1, so that coarse = (s(2xc-1) + 2 s(2xc) + s(2xc+1)) / 64:
    // MGkernelRestrict3D's 27 terms and weights, summed in FP64 in another
    // order (the same FP32 / FP16 results, FP64 ones to the last bit)
    constexpr bool rst = XT == 2 && MODE == MARCH_JACOBI && !mixed;
    static_assert(!rst || V * P <= 8, "the restriction holds two coarse points per thread");
    T* rw          = pl0 + NB * HW;
    const int ec   = e >> 1;
    double racc[2] = { 0., 0. };
    auto rplane    = [&](int X) {
        const T* R = rw + (X & 1) * E2;
        const int64_t co = a.offc[j] + (int64_t)(X >> 1) * ec * ec;
        T* dst           = reinterpret_cast<T*>(a.rc) + co;
#pragma unroll
        for (int u = 0; u < 2; u++)
        {
            const int cp = threadIdx.x + u * NT;
            if (cp >= ec * ec) break;
            const int yc = cp / ec, zc = cp - (cp / ec) * ec;
            const T* r0  = R + 2 * yc * e + 2 * zc;
            auto row     = [&](const T* rr) -> double {
                double t = 2. * (double)rr[0] + (double)rr[1];
                if (zc > 0) t += (double)rr[-1];
                return t;
            };
            double sp = 2. * row(r0) + row(r0 + e);
            if (yc > 0) sp += row(r0 - e);
            if (X & 1)
            {
                racc[u] += sp;
                const T cv = (T)(kInv64 * racc[u]);
                dst[cp]    = cv;
                if (a.rx)
                {
                    T xv = (T)0;
                    xv += static_cast<T>(a.rjf * static_cast<double>(cv));
                    reinterpret_cast<T*>(a.rx)[co + cp] = xv;
                }
                racc[u] = sp;
            }
            else
                racc[u] += 2. * sp;
        }
    };
    // the fused prolongation: coarse plane k (zero for k >= ec) is loaded to
    // registers and stored to ring slot k % 3 (row stride ec + 1); plane
    // m + 3 is staged in sweep iteration 2m, two iterations before the first
    // fine plane that reads it is corrected (plane 2m + 5 at the end of
    // iteration 2m + 2), and after the last read of plane m (fine plane
    // 2m + 1, iteration 2m - 2)
    static_assert(!pro || V * P <= 8, "the prolongation stages two coarse points per thread");
    const T* pcs = reinterpret_cast<const T*>(a.pc);
    T* cring     = pl0 + NB * HW;
    auto cload   = [&](int k, T (&cv)[2]) {
#pragma unroll
        for (int u = 0; u < 2; u++)
        {
            const int cp = threadIdx.x + u * NT;
            cv[u]        = (k < ec && cp < ec * ec) ? pcs[a.offc[j] + (int64_t)k * ec * ec + cp] : (T)0;
        }
    };
    auto cstore = [&](int k, const T (&cv)[2]) {
        T* sl = cring + (k % 3) * a.cpl;
#pragma unroll
        for (int u = 0; u < 2; u++)
        {
            const int cp = threadIdx.x + u * NT;
            if (cp < ec * ec) sl[(cp / ec) * (ec + 1) + cp - (cp / ec) * ec] = cv[u];
        }
    };
    // lane group k of fine plane pn: val -= P(coarse), prolong8_val on the
    // coarse values (xc + a, yc + b, zc + d) of each point, as k_box_prolong8
    auto pcorr = [&](int pn, V4& val, int k) {
        const int xc = pn >> 1, yc = qy[k] >> 1, zb = qz[k] >> 1;
        const T* s0  = cring + (xc % 3) * a.cpl + yc * (ec + 1) + zb;
        const T* s1  = cring + ((xc + 1) % 3) * a.cpl + yc * (ec + 1) + zb;
        T cz[2][2][V / 2 + 1];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int d = 0; d <= V / 2; d++)
            {
                cz[0][b][d] = s0[b * (ec + 1) + d];
                cz[1][b][d] = s1[b * (ec + 1) + d];
            }
#pragma unroll
        for (int i = 0; i < V; i++)
        {
            T c[2][2][2];
#pragma unroll
            for (int aa = 0; aa < 2; aa++)
#pragma unroll
                for (int b = 0; b < 2; b++)
#pragma unroll
                    for (int d = 0; d < 2; d++)
                        c[aa][b][d] = cz[aa][b][(i >> 1) + d];
            const T pv = prolong8_val(c, pn & 1, qy[k] & 1, i & 1);
            T v        = val[i];
            v += static_cast<T>(-1. * static_cast<double>(pv));
            val[i] = v;
        }
    };
#pragma unroll
    for (int k = 0; k < P; k++)
    {
        const int q = V * (threadIdx.x + k * NT);
        qv[k][0]    = zero4;
        qv[k][1]    = zero4;
        qv[k][2]    = act[k] ? ld(base + q) : zero4;
        qv[k][3]    = (act[k] && e > 1) ? ld(base + E2 + q) : zero4;
        qv[k][4]    = (act[k] && e > 2) ? ld(base + 2 * E2 + q) : zero4;
    }
    if constexpr (pro)
    {
        T cv[2];
        __syncthreads();  // the ring's zero fill before its first planes
        for (int kc = 0; kc < 3; kc++)
        {
            cload(kc, cv);
            cstore(kc, cv);
        }
    }
    __syncthreads();
    if constexpr (pro)
    {
#pragma unroll
        for (int k = 0; k < P; k++)
            if (act[k])
            {
                pcorr(0, qv[k][2], k);
                pcorr(1, qv[k][3], k);
                pcorr(2, qv[k][4], k);
            }
    }
    for (int x = 0; x < e; x++)
    {
        T* pl = (NB == 2 && (x & 1)) ? pl0 + HW : pl0;
        double pot[P][V];
        T cv[2];
        if (pro && !(x & 1)) cload((x >> 1) + 3, cv);
#pragma unroll
        for (int k = 0; k < P; k++)
        {
            const int q = V * (threadIdx.x + k * NT);
            nx[k]       = (act[k] && x + 3 < e) ? ld(base + (int64_t)(x + 3) * E2 + q) : zero4;
#pragma unroll
            for (int i = 0; i < V; i++)
                pot[k][i] = 0.;
            if (MODE == MARCH_HLOC && TYPE == MGX_LAP_4TH && act[k])
            {
                int gx = a.origin[3 * j] + x, gy = a.origin[3 * j + 1] + qy[k];
                if (gx >= a.g0) gx -= a.g0;
                if (gy >= a.g1) gy -= a.g1;
                const double* vrow = vtot + ((int64_t)gx * a.g1 + gy) * a.g2;
#pragma unroll
                for (int i = 0; i < V; i++)
                {
                    int gz = a.origin[3 * j + 2] + qz[k] + i;
                    if (gz >= a.g2) gz -= a.g2;
                    pot[k][i] = vrow[gz];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < P; k++)
            if (act[k]) *reinterpret_cast<V4*>(pl + (qy[k] + 2) * W + qz[k] + ZR) = qv[k][2];
        __syncthreads();
        if (rst && x > 0) rplane(x - 1);
        if (pro && !(x & 1)) cstore((x >> 1) + 3, cv);
#pragma unroll
        for (int k = 0; k < P; k++)
        {
            if (!act[k]) continue;
            const int q      = V * (threadIdx.x + k * NT);
            const T* ctr     = pl + (qy[k] + 2) * W + qz[k] + ZR;
            const int64_t gi = base + (int64_t)x * E2 + q;
            const V4 ym1 = *reinterpret_cast<const V4*>(ctr - W);
            const V4 yp1 = *reinterpret_cast<const V4*>(ctr + W);
            const V2 zl  = *reinterpret_cast<const V2*>(ctr - 2);
            const V2 zr  = *reinterpret_cast<const V2*>(ctr + V);
            V4 ym2 = zero4, yp2 = zero4;
            if (TYPE == MGX_LAP_4TH)
            {
                ym2 = *reinterpret_cast<const V4*>(ctr - 2 * W);
                yp2 = *reinterpret_cast<const V4*>(ctr + 2 * W);
            }
            const V4 cc = qv[k][2];
            T zz[V + 4];
            zz[0] = zl[0];
            zz[1] = zl[1];
#pragma unroll
            for (int i = 0; i < V; i++)
                zz[2 + i] = cc[i];
            zz[V + 2] = zr[0];
            zz[V + 3] = zr[1];
            V4 fv = zero4;
            if (MODE == MARCH_JACOBI) fv = *reinterpret_cast<const V4*>(f + gi);
            V4 ov, wv4;
#pragma unroll
            for (int i = 0; i < V; i++)
            {
                // FP64 arithmetic on FP32 / FP64 values as the reference; FP32
                // on FP16 values (the FP16 V-cycle is a preconditioner of
                // 11-bit values: FP64 would only cost issue slots)
                using C = typename std::conditional<sizeof(T) == 2, float, double>::type;
                C r;
                if (TYPE == MGX_LAP_4TH)
                {
                    C diag = (C)c4.c0;
                    if (MODE == MARCH_HLOC) diag = (C)(c4.c0 + pot[k][i]);
                    r = diag * (C)cc[i] + (C)c4.c1x * ((C)qv[k][1][i] + (C)qv[k][3][i])
                        + (C)c4.c1y * ((C)ym1[i] + (C)yp1[i]) + (C)c4.c1z * ((C)zz[i + 1] + (C)zz[i + 3])
                        + (C)c4.c2x * ((C)qv[k][0][i] + (C)qv[k][4][i])
                        + (C)c4.c2y * ((C)ym2[i] + (C)yp2[i]) + (C)c4.c2z * ((C)zz[i] + (C)zz[i + 4]);
                }
                else if (sizeof(T) == 2)
                {
                    r = (C)c2.c0 * (C)cc[i] + (C)c2.c1x * ((C)qv[k][1][i] + (C)qv[k][3][i])
                        + (C)c2.c1y * ((C)ym1[i] + (C)yp1[i]) + (C)c2.c1z * ((C)zz[i + 1] + (C)zz[i + 3]);
                }
                else
                {
                    r = c2.c0 * cc[i] + c2.c1x * (T)(qv[k][1][i] + qv[k][3][i]) + c2.c1y * (T)(ym1[i] + yp1[i])
                        + c2.c1z * (T)(zz[i + 1] + zz[i + 3]);
                }
                if (MODE == MARCH_JACOBI)
                {
                    T wv = (T)r;
                    wv += static_cast<T>((C)-1. * static_cast<C>(fv[i]));
                    wv4[i] = wv;
                    T xv   = cc[i];
                    xv += static_cast<T>((C)a.mjf * static_cast<C>(wv));
                    ov[i] = xv;
                }
                else
                    ov[i] = (T)r;
            }
            if (MODE == MARCH_JACOBI && wout) *reinterpret_cast<V4*>(wout + gi) = wv4;
            if (rst) *reinterpret_cast<V4*>(rw + (x & 1) * E2 + q) = wv4;
            VO oo;
            if constexpr (mixed)
            {
#pragma unroll
                for (int i = 0; i < V; i++)
                    oo[i] = (TO)((double)ov[i] * oj);
                *reinterpret_cast<VO*>(out + gi) = oo;
            }
            else
            {
                oo = ov;
                *reinterpret_cast<V4*>(out + gi) = ov;
            }
            if (dots)
            {
                const VO k0 = *reinterpret_cast<const VO*>(dkr0 + gi);
#pragma unroll
                for (int i = 0; i < V; i++)
                {
                    const double fa = a.da * (double)oo[i], df = fa - a.da * (double)k0[i];
                    dacc0 += df * df;
                    dacc1 += df * fa;
                }
            }
        }
        if (NB == 1) __syncthreads();
#pragma unroll
        for (int k = 0; k < P; k++)
        {
            if (pro && act[k] && x + 3 < e) pcorr(x + 3, nx[k], k);
            qv[k][0] = qv[k][1];
            qv[k][1] = qv[k][2];
            qv[k][2] = qv[k][3];
            qv[k][3] = qv[k][4];
            qv[k][4] = nx[k];
        }
    }
    if (rst)
    {
        __syncthreads();
        rplane(e - 1);
    }
    if (dots)
    {
        __shared__ double dred[2][16];
        for (int o = 32; o > 0; o >>= 1)
        {
            dacc0 += __shfl_xor(dacc0, o, 64);
            dacc1 += __shfl_xor(dacc1, o, 64);
        }
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0)
        {
            dred[0][wv] = dacc0;
            dred[1][wv] = dacc1;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double s0 = 0., s1 = 0.;
            for (int w = 0; w < NT / 64; w++)
            {
                s0 += dred[0][w];
                s1 += dred[1][w];
            }
            a.dout[j]             = s0 * a.dvel;
            a.dout[a.nbox + j]    = s1 * a.dvel;
        }
    }
}

// MGkernelRestrict3D on a box (fine values outside the box are zero)
template <typename T>
__global__ void __launch_bounds__(256) k_box_restrict2(const T* fine, const int64_t* off_f, T* coarse, BoxView vc)
{
    int j, ec, x, y, z, p;
    int64_t cbase;
    if (!tile_point(vc, &j, &ec, &x, &y, &z, &cbase, &p)) return;
    const int ef = 2 * ec;
    const T* u   = fine + off_f[j];
    const int X = 2 * x, Y = 2 * y, Z = 2 * z;
#define F(dx, dy, dz) (double)bv(u, ef, X + (dx), Y + (dy), Z + (dz))
    double face   = F(1, 0, 0) + F(-1, 0, 0) + F(0, 1, 0) + F(0, -1, 0) + F(0, 0, -1) + F(0, 0, 1);
    double corner = F(1, 1, -1) + F(1, 1, 1) + F(1, -1, -1) + F(1, -1, 1) + F(-1, 1, -1) + F(-1, 1, 1)
                    + F(-1, -1, -1) + F(-1, -1, 1);
    double edge = F(0, 1, -1) + F(0, 1, 1) + F(0, -1, -1) + F(0, -1, 1) + F(1, 0, -1) + F(1, 0, 1) + F(-1, 0, -1)
                  + F(-1, 0, 1) + F(-1, -1, 0) + F(1, -1, 0) + F(-1, 1, 0) + F(1, 1, 0);
    coarse[cbase + p] = (T)(kInv64 * (8. * F(0, 0, 0) + 4. * face + 2. * edge + corner));
#undef F
}

// k_box_restrict2 for two z-adjacent coarse points per thread (every coarse
// edge even, so a pair never leaves its row): the fine values z = 2z-1 ..
// 2z+3 of the 9 (x, y) rows around them, the four from 2z in vector loads;
// per coarse point the same double sums in the same order as k_box_restrict2
template <typename T>
__global__ void __launch_bounds__(kTile / 2) k_box_restrict2p(const T* fine, const int64_t* off_f, T* coarse,
    BoxView vc)
{
    // loads of four values (16 bytes FP32, 8 bytes FP16) or two (FP64)
    constexpr int VN = sizeof(T) == 8 ? 2 : 4;
    typedef T vec16 __attribute__((ext_vector_type(VN)));
    const int64_t t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= vc.ntiles) return;
    const int j  = vc.tile[2 * t];
    const int ec = vc.edge[j] >> vc.level;
    const int p  = vc.tile[2 * t + 1] + 2 * (int)threadIdx.x;
    if (p >= ec * ec * ec) return;
    int r       = p;
    const int x = r / (ec * ec);
    r -= x * ec * ec;
    const int y = r / ec;
    const int z = r - y * ec;
    const int ef = 2 * ec;
    const T* u   = fine + off_f[j];
    const int X = 2 * x, Y = 2 * y, Z = 2 * z;
    // w[dx + 1][dy + 1][k] = fine(X + dx, Y + dy, Z - 1 + k), zero outside the box
    T w[3][3][5];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
        {
            const int xx = X + a - 1, yy = Y + b - 1;
            if ((unsigned)xx >= (unsigned)ef || (unsigned)yy >= (unsigned)ef)
            {
                for (int k = 0; k < 5; k++)
                    w[a][b][k] = (T)0;
                continue;
            }
            const T* row = u + ((int64_t)xx * ef + yy) * ef;
            w[a][b][0]   = Z > 0 ? row[Z - 1] : (T)0;
#pragma unroll
            for (int q = 0; q < 4; q += VN)
            {
                const vec16 l = *reinterpret_cast<const vec16*>(row + Z + q);
#pragma unroll
                for (int i = 0; i < VN; i++)
                    w[a][b][1 + q + i] = l[i];
            }
        }
    const int64_t cbase = vc.off[j];
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
#define F(dx, dy, dz) (double)w[(dx) + 1][(dy) + 1][2 * h + 1 + (dz)]
        double face = F(1, 0, 0) + F(-1, 0, 0) + F(0, 1, 0) + F(0, -1, 0) + F(0, 0, -1) + F(0, 0, 1);
        double corner = F(1, 1, -1) + F(1, 1, 1) + F(1, -1, -1) + F(1, -1, 1) + F(-1, 1, -1) + F(-1, 1, 1)
                        + F(-1, -1, -1) + F(-1, -1, 1);
        double edge = F(0, 1, -1) + F(0, 1, 1) + F(0, -1, -1) + F(0, -1, 1) + F(1, 0, -1) + F(1, 0, 1)
                      + F(-1, 0, -1) + F(-1, 0, 1) + F(-1, -1, 0) + F(1, -1, 0) + F(-1, 1, 0) + F(1, 1, 0);
        coarse[cbase + p + h] = (T)(kInv64 * (8. * F(0, 0, 0) + 4. * face + 2. * edge + corner));
#undef F
    }
}

// MGkernelExtend3D value at a fine box point (injected coarse values, zero
// outside the coarse box)
template <typename T>
__device__ __forceinline__ T box_prolong_val(const T* uc, int ec, int x, int y, int z)
{
    const int ox = x & 1, oy = y & 1, oz = z & 1;
#define I(dx, dy, dz) bv(uc, ec, (x + (dx)) >> 1, (y + (dy)) >> 1, (z + (dz)) >> 1)
    if (!ox && !oy && !oz) return I(0, 0, 0);
    if (!ox && !oy && oz) return (T)(0.5 * (T)(I(0, 0, -1) + I(0, 0, 1)));
    if (!ox && oy && !oz) return (T)(0.5 * (T)(I(0, 1, 0) + I(0, -1, 0)));
    if (!ox && oy && oz) return (T)(0.25 * (T)(I(0, 1, 1) + I(0, -1, 1) + I(0, 1, -1) + I(0, -1, -1)));
    if (ox && !oy && oz) return (T)(0.25 * (T)(I(1, 0, 1) + I(1, 0, -1) + I(-1, 0, 1) + I(-1, 0, -1)));
    if (ox && !oy && !oz) return (T)(0.5 * (T)(I(1, 0, 0) + I(-1, 0, 0)));
    if (ox && oy && oz)
        return (T)(0.125
                   * (T)(I(1, 1, 1) + I(1, 1, -1) + I(1, -1, 1) + I(1, -1, -1) + I(-1, 1, 1) + I(-1, 1, -1)
                         + I(-1, -1, 1) + I(-1, -1, -1)));
    return (T)(0.25 * (T)(I(1, 1, 0) + I(1, -1, 0) + I(-1, 1, 0) + I(-1, -1, 0)));
#undef I
}

// The 8 fine values of the coarse cell at (xc, yc, zc): fine point
// (2xc+ox, 2yc+oy, 2zc+oz) reads coarse indices xc (and xc+1 when ox = 1),
// so with c[a][b][d] = coarse(xc+a, yc+b, zc+d) the same expressions as
// box_prolong_val apply with I(dx, dy, dz) = c[dx>0][dy>0][dz>0].
template <typename T>
__device__ __forceinline__ T prolong8_val(const T (&c)[2][2][2], int ox, int oy, int oz)
{
#define I(dx, dy, dz) c[(dx) > 0][(dy) > 0][(dz) > 0]
    if (!ox && !oy && !oz) return I(0, 0, 0);
    if (!ox && !oy && oz) return (T)(0.5 * (T)(I(0, 0, -1) + I(0, 0, 1)));
    if (!ox && oy && !oz) return (T)(0.5 * (T)(I(0, 1, 0) + I(0, -1, 0)));
    if (!ox && oy && oz) return (T)(0.25 * (T)(I(0, 1, 1) + I(0, -1, 1) + I(0, 1, -1) + I(0, -1, -1)));
    if (ox && !oy && oz) return (T)(0.25 * (T)(I(1, 0, 1) + I(1, 0, -1) + I(-1, 0, 1) + I(-1, 0, -1)));
    if (ox && !oy && !oz) return (T)(0.5 * (T)(I(1, 0, 0) + I(-1, 0, 0)));
    if (ox && oy && oz)
        return (T)(0.125
                   * (T)(I(1, 1, 1) + I(1, 1, -1) + I(1, -1, 1) + I(1, -1, -1) + I(-1, 1, 1) + I(-1, 1, -1)
                         + I(-1, -1, 1) + I(-1, -1, -1)));
    return (T)(0.25 * (T)(I(1, 1, 0) + I(1, -1, 0) + I(-1, 1, 0) + I(-1, -1, 0)));
#undef I
}

// prolongation by coarse cells: one thread per coarse point writes its 8
// fine points (no parity divergence, each coarse value loaded once per
// cell); with SUB: v -= P(coarse)
template <typename T, bool SUB>
__global__ void __launch_bounds__(256) k_box_prolong8(const T* coarse, const int64_t* off_f, T* fine, BoxView vc)
{
    int j, ec, xc, yc, zc, p;
    int64_t cbase;
    if (!tile_point(vc, &j, &ec, &xc, &yc, &zc, &cbase, &p)) return;
    const T* uc = coarse + cbase;
    T c[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int d = 0; d < 2; d++)
                c[a][b][d] = bv(uc, ec, xc + a, yc + b, zc + d);
    const int ef      = 2 * ec;
    T* uf             = fine + off_f[j];
#pragma unroll
    for (int ox = 0; ox < 2; ox++)
#pragma unroll
        for (int oy = 0; oy < 2; oy++)
        {
            const int64_t row = ((int64_t)(2 * xc + ox) * ef + (2 * yc + oy)) * ef + 2 * zc;
#pragma unroll
            for (int oz = 0; oz < 2; oz++)
            {
                const T pv = prolong8_val(c, ox, oy, oz);
                if (SUB)
                {
                    T v = uf[row + oz];
                    v += static_cast<T>(-1. * static_cast<double>(pv));
                    uf[row + oz] = v;
                }
                else
                    uf[row + oz] = pv;
            }
        }
}

// prolongation; with SUB: v -= P(coarse) (extend3D, app_mask, operator-=)
template <typename T, bool SUB>
__global__ void __launch_bounds__(256) k_box_prolong(const T* coarse, const int64_t* off_c, T* fine, BoxView vf)
{
    int j, ef, x, y, z, p;
    int64_t fbase;
    if (!tile_point(vf, &j, &ef, &x, &y, &z, &fbase, &p)) return;
    const T pv = box_prolong_val(coarse + off_c[j], ef >> 1, x, y, z);
    if (SUB)
    {
        T v = fine[fbase + p];
        v += static_cast<T>(-1. * static_cast<double>(pv));
        fine[fbase + p] = v;
    }
    else
        fine[fbase + p] = pv;
}

template <typename T>
__global__ void k_fill_b(T* u, int64_t n, T v)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        u[t] = v;
}

// precond_mg entry: f = (T)res; x = 0 + (T)(gamma*(double)f)
template <typename T>
__global__ void k_prep(const double* res, T* f, T* x, int64_t n, double gamma)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    {
        const T fv = (T)res[t];
        f[t]       = fv;
        T xv       = (T)0;
        xv += static_cast<T>(gamma * static_cast<double>(fv));
        x[t] = xv;
    }
}

// The first Jacobi sweep from x = 0, done pointwise: w = A*0 - f = -f, so
// x = 0 + (T)(-jf * w) = 0 + (T)(jf * f), the value GridFuncVector::jacobi
// produces (the mask is 1 on the box).  With res non-null it also forms
// f = (T)res (the fine level's FP64 -> T conversion).
template <typename T>
__global__ void k_first_sweep(const double* res, T* f, T* x, int64_t n, double jf)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    {
        T fv;
        if (res)
        {
            fv   = (T)res[t];
            f[t] = fv;
        }
        else
            fv = f[t];
        T xv = (T)0;
        xv += static_cast<T>(jf * static_cast<double>(fv));
        x[t] = xv;
    }
}

// per-box scale of the residual for the FP16 V-cycle: sc_j brings the box's
// RMS of R to kF16Rms (FP16 keeps 11 bits between 6e-5 and 65504; the
// V-cycle is linear, so its FP32 output is divided by sc_j); inv_j = 1/sc_j.
// dres_j = vel |R_j|^2 (may alias inv)
constexpr double kF16Rms = 8.;
__global__ void k_res_scale(const double* dres, const int64_t* off, int nbox, double vel, double* sc, double* inv)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nbox) return;
    const double ms = dres[j] / (vel * (double)(off[j + 1] - off[j]));
    const double c  = (ms > 0. && ms < 1e300) ? kF16Rms / sqrt(ms) : 1.;
    sc[j]           = c;
    inv[j]          = 1. / c;
}

// k_first_sweep / k_prep with the box's scale: f = (T)(sc_j res); the zero
// start writes x0 = 0 + (T)(c f) with c = the Jacobi factor, the gamma start
// c = gamma.  One block per box, four points per thread and step (box sizes
// multiples of 4)
template <typename T>
__global__ void __launch_bounds__(256) k_first_sweep_box(const double* res, const double* sc, T* f, T* x, double c,
    const int64_t* off, int nbox)
{
    typedef double d4 __attribute__((ext_vector_type(4)));
    typedef T t4 __attribute__((ext_vector_type(4)));
    const int j = blockIdx.x;
    if (j >= nbox) return;
    const double s   = sc[j];
    const int64_t q0 = off[j], q1 = off[j + 1];
    for (int64_t q = q0 + 4 * (int64_t)threadIdx.x; q < q1; q += 4 * (int64_t)blockDim.x)
    {
        const d4 r = *reinterpret_cast<const d4*>(res + q);
        t4 fv, xv;
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            fv[i]  = (T)(s * r[i]);
            T xi   = (T)0;
            xi += static_cast<T>(c * static_cast<double>(fv[i]));
            xv[i] = xi;
        }
        *reinterpret_cast<t4*>(f + q) = fv;
        *reinterpret_cast<t4*>(x + q) = xv;
    }
}

template <typename T>
__global__ void k_to_f64(const T* x, double* out, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (double)x[t];
}

// psi += alpha * (double)x  (ABPG::update_states, src/ABPG.cc:110,129-130)
template <typename T>
__global__ void k_update(double* psi, const T* x, int64_t n, double alpha)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        psi[t] += alpha * (double)x[t];
}

// -------------------------------------------------------- reductions -----
__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++)
    {
        v[k] = wave_sum(v[k]);
        if (lane == 0) lds[wid * NV + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; k++)
        {
            double s = 0.;
            for (int w = 0; w < nw; w++)
                s += lds[w * NV + k];
            v[k] = s;
        }
    }
}

// Stencil part of (H psi_B)(X) at a point X outside box B along axis d only
// (layer 1 or 2 beyond face `side`), from the legs that land inside B.
__device__ __forceinline__ double skin_hv(const double* pB, int eB, int d, int side, int layer, int X0, int X1,
    int X2, double c1, double c2)
{
    // X0..X2: coordinates, the one along d is outside
    int q[3] = { X0, X1, X2 };
    auto at  = [&](int v) {
        q[d] = v;
        return v >= 0 && v < eB ? pB[((int64_t)q[0] * eB + q[1]) * eB + q[2]] : 0.;  // edge-1 boxes
    };
    if (side == 0)
        return layer == 1 ? c1 * at(0) + c2 * at(1) : c2 * at(0);
    return layer == 1 ? c1 * at(eB - 1) + c2 * at(eB - 2) : c2 * at(eB - 1);
}

// sum over r in box A, r outside box B (within reach 2), psi_A(r) (H psi_B)(r);
// shift maps A-local -> B-local coordinates.  Threads of the block stride.
__device__ double skin_part(const double* pA, int eA, const double* pB, int eB, const int sh[3], const Coef4& c)
{
    double acc = 0.;
    const double c1[3] = { c.c1x, c.c1y, c.c1z }, c2[3] = { c.c2x, c.c2y, c.c2z };
    int lo[3], hi[3];
    for (int k = 0; k < 3; k++)
    {
        lo[k] = max(0, sh[k]);
        hi[k] = min(eB, sh[k] + eA);
    }
    for (int slab = 0; slab < 12; slab++)
    {
        const int d = slab >> 2, side = (slab >> 1) & 1, layer = (slab & 1) + 1;
        const int Xd = side == 0 ? -layer : eB - 1 + layer;
        const int xd = Xd - sh[d];
        if (xd < 0 || xd >= eA) continue;
        const int u = (d + 1) % 3, w = (d + 2) % 3;
        const int nu = hi[u] - lo[u], nw = hi[w] - lo[w];
        if (nu <= 0 || nw <= 0) continue;
        for (int t = threadIdx.x; t < nu * nw; t += blockDim.x)
        {
            const int a = t / nw, b = t - (t / nw) * nw;
            int X[3];
            X[d] = Xd;
            X[u] = lo[u] + a;
            X[w] = lo[w] + b;
            const double hv = skin_hv(pB, eB, d, side, layer, X[0], X[1], X[2], c1[d], c2[d]);
            const int x0 = X[0] - sh[0], x1 = X[1] - sh[1], x2 = X[2] - sh[2];
            acc += pA[((int64_t)x0 * eA + x1) * eA + x2] * hv;
        }
    }
    return acc;
}

// S / H partial sums of one interacting pair (one block per pair):
// overlap term from the stored (masked) H psi, plus the stencil reach
// outside each box, so that H_ij = <psi_i, H psi_j> (Galerkin).
template <int PREC>
__global__ void __launch_bounds__(256) k_pair_gram(const double* psi, const double* hpsi, const int64_t* off,
    const int* edge, const int* pairs, int64_t npairs, Coef4 c4, double* out)
{
    __shared__ double lds[3 * 4];
    const int64_t pid = blockIdx.x;
    if (pid >= npairs) return;
    const int* P = pairs + 11 * pid;
    const int i = P[0], j = P[1];
    const int a0 = P[2], a1 = P[3], a2 = P[4], b0 = P[5], b1 = P[6], b2 = P[7];
    const int s0 = P[8], s1 = P[9], s2 = P[10];
    const int ei = edge[i], ej = edge[j];
    const int n1 = max(b1 - a1, 0), n2 = max(b2 - a2, 0);
    const int64_t npt = (int64_t)max(b0 - a0, 0) * n1 * n2;
    const double* pi  = psi + off[i];
    const double* pj  = psi + off[j];
    const double* hi  = hpsi + off[i];
    const double* hj  = hpsi + off[j];
    double acc[3] = { 0., 0., 0. };
    float fs = 0.f, fh1 = 0.f, fh2 = 0.f, fx = 0.f, fx1 = 0.f, fx2 = 0.f;
    int cnt = 0;
    for (int64_t q = threadIdx.x; q < npt; q += blockDim.x)
    {
        int r       = (int)q;
        const int x = r / (n1 * n2);
        r -= x * n1 * n2;
        const int y = r / n2;
        const int z = r - y * n2;
        const int X = a0 + x, Y = a1 + y, Z = a2 + z;
        const int64_t ii = ((int64_t)X * ei + Y) * ei + Z;
        const int64_t jj = ((int64_t)(X + s0) * ej + (Y + s1)) * ej + (Z + s2);
        const double vi = pi[ii], vj = pj[jj], wi = hi[ii], wj = hj[jj];
        if (PREC == MGX_PREC_F64)
        {
            acc[0] += vi * vj;
            acc[1] += vi * wj;
            acc[2] += wi * vj;
        }
        else
        {
            // hi/lo split: x = hi + lo, hi = fl32(x), lo = fl32(x - hi);
            // products hi*hi + (hi*lo + lo*hi) in FP32, chunks summed in FP64
            const float vih = (float)vi, vil = (float)(vi - (double)vih);
            const float vjh = (float)vj, vjl = (float)(vj - (double)vjh);
            const float wih = (float)wi, wil = (float)(wi - (double)wih);
            const float wjh = (float)wj, wjl = (float)(wj - (double)wjh);
            fs += vih * vjh;
            fx += vih * vjl + vil * vjh;
            fh1 += vih * wjh;
            fx1 += vih * wjl + vil * wjh;
            fh2 += wih * vjh;
            fx2 += wih * vjl + wil * vjh;
            if (++cnt == 16)
            {
                acc[0] += (double)fs + (double)fx;
                acc[1] += (double)fh1 + (double)fx1;
                acc[2] += (double)fh2 + (double)fx2;
                fs = fh1 = fh2 = fx = fx1 = fx2 = 0.f;
                cnt = 0;
            }
        }
    }
    if (PREC != MGX_PREC_F64)
    {
        acc[0] += (double)fs + (double)fx;
        acc[1] += (double)fh1 + (double)fx1;
        acc[2] += (double)fh2 + (double)fx2;
    }
    const int shij[3] = { s0, s1, s2 }, shji[3] = { -s0, -s1, -s2 };
    acc[1] += skin_part(pi, ei, pj, ej, shij, c4);
    acc[2] += skin_part(pj, ej, pi, ei, shji, c4);
    block_sum<3>(acc, lds);
    if (threadIdx.x == 0)
    {
        out[3 * pid]     = acc[0];
        out[3 * pid + 1] = acc[1];
        out[3 * pid + 2] = acc[2];
    }
}

// ------------------------------------------------ Mehrstellen on boxes ----
// Hamiltonian<T>::applyLocal, Mehrstellen branch (src/Hamiltonian.cc:106-131),
// on a box (psi zero outside): w1 = psi V, work1 = B w1, w1 = -Lap_M psi,
// work1 += w1; the arithmetic of k_hlocM_packed with the box's zero ring in
// place of the periodic wrap.  V is gathered from the periodic grid.
__global__ void __launch_bounds__(256) k_box_hlocM(const double* in, const double* vtot, double* out, BoxView v,
    CoefM cm)
{
    int j, e, x, y, z, p;
    int64_t base;
    if (!tile_point(v, &j, &e, &x, &y, &z, &base, &p)) return;
    const double* a = in + base;
    const int ox = v.origin[3 * j], oy = v.origin[3 * j + 1], oz = v.origin[3 * j + 2];
    auto vpsi = [&](int dx, int dy, int dz) -> double {
        const int xx = x + dx, yy = y + dy, zz = z + dz;
        if ((unsigned)xx >= (unsigned)e || (unsigned)yy >= (unsigned)e || (unsigned)zz >= (unsigned)e) return 0.;
        int gx = ox + xx, gy = oy + yy, gz = oz + zz;
        if (gx >= v.g0) gx -= v.g0;
        if (gy >= v.g1) gy -= v.g1;
        if (gz >= v.g2) gz -= v.g2;
        return a[((int64_t)xx * e + yy) * e + zz] * vtot[((int64_t)gx * v.g1 + gy) * v.g2 + gz];
    };
    const double bw = 0.5 * vpsi(0, 0, 0)
                      + kInv12 * (vpsi(-1, 0, 0) + vpsi(1, 0, 0) + vpsi(0, -1, 0) + vpsi(0, 1, 0) + vpsi(0, 0, -1)
                                     + vpsi(0, 0, 1));
    Coef4 c4{};
    Coef2 c2{};
    const double lm = box_stencil<double, MGX_LAP_MEHRSTELLEN>(a, e, x, y, z, c4, c2, cm, 0.);
    double o = bw;
    o += 1. * lm;
    out[base + p] = o;
}

// Lap::rhs (FDkernelRHS_4th_Mehr1, src/pb/FDkernels.cc:523-584) on a box, zero
// outside: the arithmetic of k_rhsM_packed
__global__ void __launch_bounds__(256) k_box_bpsiM(const double* in, double* out, BoxView v)
{
    int j, e, x, y, z, p;
    int64_t base;
    if (!tile_point(v, &j, &e, &x, &y, &z, &base, &p)) return;
    const double* a = in + base;
#define A(dx, dy, dz) bv(a, e, x + (dx), y + (dy), z + (dz))
    out[base + p] = 0.5 * A(0, 0, 0)
                    + kInv12 * (A(-1, 0, 0) + A(1, 0, 0) + A(0, -1, 0) + A(0, 1, 0) + A(0, 0, -1) + A(0, 0, 1));
#undef A
}

// The reach of the Mehrstellen H psi_B and B psi_B at X (B-local), one point
// outside box B in one axis (a face: axis coefficient, the edge neighbours on
// the face, (1/12) V psi of B(V psi)) or two axes (an edge: the edge
// coefficient); q is the box point next to X.
struct MehrReach
{
    double h, b;
};

__device__ __forceinline__ double coef_axis(const CoefM& c, int d) { return d == 0 ? c.cx : (d == 1 ? c.cy : c.cz); }

__device__ __forceinline__ double coef_edge(const CoefM& c, int d, int u)
{
    const int m = d + u;  // 1: xy, 3: yz, 2: xz
    return m == 1 ? c.cxy : (m == 3 ? c.cyz : c.cxz);
}

// sum over the points x of box A outside box B within its one-point reach of
// psi_A(x) (H psi_B)(x) and psi_A(x) (B psi_B)(x); sh maps A-local -> B-local;
// block-stride loops over the faces and edges of B
__device__ void mehr_skin(const double* pA, int eA, const double* pB, int eB, const int* oB, const int sh[3],
    const double* vtot, int g0, int g1, int g2, const CoefM& cm, double* hs, double* bs)
{
    // A's extent in B-local coordinates: [sh, sh + eA)
    int lo[3], hi[3];
    for (int k = 0; k < 3; k++)
    {
        lo[k] = sh[k];
        hi[k] = sh[k] + eA;
    }
    auto pa = [&](int X0, int X1, int X2) -> double {
        return pA[((int64_t)(X0 - sh[0]) * eA + (X1 - sh[1])) * eA + (X2 - sh[2])];
    };
    auto pb = [&](int q0, int q1, int q2) -> double { return pB[((int64_t)q0 * eB + q1) * eB + q2]; };
    auto inB = [&](int q) -> bool { return q >= 0 && q < eB; };
    const int gd[3] = { g0, g1, g2 };
    auto vB = [&](const int q[3]) -> double {
        int g[3];
        for (int k = 0; k < 3; k++)
        {
            g[k] = oB[k] + q[k];
            if (g[k] >= gd[k]) g[k] -= gd[k];
        }
        return vtot[((int64_t)g[0] * g1 + g[1]) * g2 + g[2]];
    };
    double h = 0., b = 0.;
    // faces: X_d = -1 or eB, the other two coordinates inside B
    for (int f = 0; f < 6; f++)
    {
        const int d = f >> 1, s = (f & 1) ? 1 : -1;
        const int Xd = s < 0 ? -1 : eB;
        if (Xd < lo[d] || Xd >= hi[d]) continue;
        const int u = (d + 1) % 3, w = (d + 2) % 3;
        const int u0 = max(0, lo[u]), u1 = min(eB, hi[u]), w0 = max(0, lo[w]), w1 = min(eB, hi[w]);
        const int nu = u1 - u0, nw = w1 - w0;
        if (nu <= 0 || nw <= 0) continue;
        for (int t = threadIdx.x; t < nu * nw; t += blockDim.x)
        {
            int X[3];
            X[d] = Xd;
            X[u] = u0 + t / nw;
            X[w] = w0 + (t - (t / nw) * nw);
            int q[3] = { X[0], X[1], X[2] };
            q[d]             = Xd - s;
            const double psq = pb(q[0], q[1], q[2]);
            double hv        = coef_axis(cm, d) * psq;
            for (int e2 : { u, w })
                for (int tt = -1; tt <= 1; tt += 2)
                {
                    int r[3] = { q[0], q[1], q[2] };
                    r[e2] += tt;
                    if (inB(r[e2])) hv += coef_edge(cm, d, e2) * pb(r[0], r[1], r[2]);
                }
            hv += kInv12 * (vB(q) * psq);
            const double av = pa(X[0], X[1], X[2]);
            h += av * hv;
            b += av * (kInv12 * psq);
        }
    }
    // edges: X_d, X_u outside by one, X_w inside
    for (int d = 0; d < 3; d++)
        for (int u = d + 1; u < 3; u++)
        {
            const int w = 3 - d - u;
            const int w0 = max(0, lo[w]), w1 = min(eB, hi[w]);
            if (w1 <= w0) continue;
            for (int c = 0; c < 4; c++)
            {
                const int sd = (c & 1) ? 1 : -1, su = (c & 2) ? 1 : -1;
                const int Xd = sd < 0 ? -1 : eB, Xu = su < 0 ? -1 : eB;
                if (Xd < lo[d] || Xd >= hi[d] || Xu < lo[u] || Xu >= hi[u]) continue;
                for (int t = threadIdx.x; t < w1 - w0; t += blockDim.x)
                {
                    int X[3];
                    X[d] = Xd;
                    X[u] = Xu;
                    X[w] = w0 + t;
                    int q[3] = { X[0], X[1], X[2] };
                    q[d] -= sd;
                    q[u] -= su;
                    h += pa(X[0], X[1], X[2]) * (coef_edge(cm, d, u) * pb(q[0], q[1], q[2]));
                }
            }
        }
    *hs += h;
    *bs += b;
}

// S, H, B-matrix partials of one interacting pair (i <= j), one block each:
// the overlap from the stored box values, plus each side's reach outside the
// other box; H and B in both orders (H is not symmetric).  part: 5 per pair
// (s, h_ij, h_ji, b_ij, b_ji).
__global__ void __launch_bounds__(256) k_pair_gram_mehr(const double* psi, const double* hpsi, const double* bpsi,
    const double* vtot, const int64_t* off, const int* edge, const int* origin, const int* pairs, int64_t npairs,
    int g0, int g1, int g2, CoefM cm, double* out)
{
    __shared__ double lds[5 * 4];
    const int64_t pid = blockIdx.x;
    if (pid >= npairs) return;
    const int* P = pairs + 11 * pid;
    const int i = P[0], j = P[1];
    const int a0 = P[2], a1 = P[3], a2 = P[4], b0 = P[5], b1 = P[6], b2 = P[7];
    const int sh[3] = { P[8], P[9], P[10] };
    const int ei = edge[i], ej = edge[j];
    const int n1 = max(b1 - a1, 0), n2 = max(b2 - a2, 0);
    const int64_t npt = (int64_t)max(b0 - a0, 0) * n1 * n2;
    const double *pi = psi + off[i], *pj = psi + off[j];
    const double *hi = hpsi + off[i], *hj = hpsi + off[j];
    const double *bi = bpsi + off[i], *bj = bpsi + off[j];
    double acc[5] = { 0., 0., 0., 0., 0. };
    for (int64_t q = threadIdx.x; q < npt; q += blockDim.x)
    {
        int r       = (int)q;
        const int x = r / (n1 * n2);
        r -= x * n1 * n2;
        const int y = r / n2;
        const int z = r - y * n2;
        const int X = a0 + x, Y = a1 + y, Z = a2 + z;
        const int64_t ii = ((int64_t)X * ei + Y) * ei + Z;
        const int64_t jj = ((int64_t)(X + sh[0]) * ej + (Y + sh[1])) * ej + (Z + sh[2]);
        const double vi = pi[ii], vj = pj[jj];
        acc[0] += vi * vj;
        acc[1] += vi * hj[jj];
        acc[2] += hi[ii] * vj;
        acc[3] += vi * bj[jj];
        acc[4] += bi[ii] * vj;
    }
    if (i != j)
    {
        const int shji[3] = { -sh[0], -sh[1], -sh[2] };
        mehr_skin(pi, ei, pj, ej, origin + 3 * j, sh, vtot, g0, g1, g2, cm, &acc[1], &acc[3]);
        mehr_skin(pj, ej, pi, ei, origin + 3 * i, shji, vtot, g0, g1, g2, cm, &acc[2], &acc[4]);
    }
    block_sum<5>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; k++)
            out[5 * pid + k] = acc[k];
}

// dense S (symmetric), H and B-matrix (column-major, entry (i, j) =
// vel <psi_i, op psi_j>) from the pair partials
__global__ void k_pair_scatter_mehr(const double* part, const int* pairs, int64_t npairs, int n, double vel,
    double* S, double* H, double* B)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < npairs; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int i = pairs[11 * t], j = pairs[11 * t + 1];
        const double* q = part + 5 * t;
        S[(int64_t)j * n + i] = q[0] * vel;
        S[(int64_t)i * n + j] = q[0] * vel;
        H[(int64_t)j * n + i] = q[1] * vel;
        H[(int64_t)i * n + j] = q[2] * vel;
        B[(int64_t)j * n + i] = q[3] * vel;
        B[(int64_t)i * n + j] = q[4] * vel;
    }
}

// R_b += (1/12) theta(a, b) psi_a(q) at the points of box b one point outside
// a face of box a (the reach of B psi_a), one block per box b over its
// interaction list (entry (a, sh): sh maps b-local -> a-local), one entry at
// a time (entries may touch the same points)
__global__ void __launch_bounds__(256) k_resid_skin_mehr(const double* psi, const double* theta, int n,
    const int64_t* off, const int* edge, const int* ilist_ptr, const int* ilist, double* res)
{
    const int b  = blockIdx.x;
    const int eb = edge[b];
    double* rb   = res + off[b];
    for (int qq = ilist_ptr[b]; qq < ilist_ptr[b + 1]; qq++)
    {
        const int* E = ilist + 4 * qq;
        const int a = E[0], ea = edge[a];
        const int sh[3] = { E[1], E[2], E[3] };
        const double t  = kInv12 * theta[(int64_t)b * n + a];
        const double* pa = psi + off[a];
        for (int f = 0; f < 6; f++)
        {
            const int d = f >> 1, s = (f & 1) ? 1 : -1;
            const int u = (d + 1) % 3, w = (d + 2) % 3;
            // the shell plane of a in b-local coordinates: Xa_d - sh_d
            const int pd = (s < 0 ? -1 : ea) - sh[d];
            if (pd < 0 || pd >= eb) continue;
            const int u0 = max(0, -sh[u]), u1 = min(eb, ea - sh[u]);
            const int w0 = max(0, -sh[w]), w1 = min(eb, ea - sh[w]);
            const int nu = u1 - u0, nw = w1 - w0;
            if (nu <= 0 || nw <= 0) continue;
            for (int k = threadIdx.x; k < nu * nw; k += blockDim.x)
            {
                int p[3];
                p[d] = pd;
                p[u] = u0 + k / nw;
                p[w] = w0 + (k - (k / nw) * nw);
                int q[3] = { p[0] + sh[0], p[1] + sh[1], p[2] + sh[2] };
                q[d] -= s;
                rb[((int64_t)p[0] * eb + p[1]) * eb + p[2]] += t * pa[((int64_t)q[0] * ea + q[1]) * ea + q[2]];
            }
        }
        __syncthreads();
    }
}

// blocks for ncell cells: a multiple of 8, so that xcd_block is a bijection
inline unsigned cell_grid(int ncell) { return (unsigned)((ncell + 7) / 8 * 8); }

constexpr int kCell  = 8;    // cell edge
constexpr int kCellK = 32;   // max boxes per cell handled by the cell kernels

typedef double d4c __attribute__((ext_vector_type(4)));
typedef float f4c __attribute__((ext_vector_type(4)));

// one box covering a cell, as the cell kernels read it: built once with the
// cell tables (base = level-0 offset of the box's data, e = its edge, b = the
// cell origin in box-local coordinates, [lo, hi) = the covered cell-local
// range), so that a block reads its entries in one dependent load
struct CellEnt
{
    int64_t base;  // offset of the box's data
    int e;
    int b0, b1, b2;
    int lo0, lo1, lo2, hi0, hi1, hi2;
    int box;
};

__device__ __forceinline__ void load_cell(const CellEnt* cce, int k, CellEnt* ce)
{
    for (int a = threadIdx.x; a < k; a += blockDim.x)
        ce[a] = cce[a];
}

// index of cell point p (x = p>>6, y = (p>>3)&7, z = p&7) in the box, or -1
__device__ __forceinline__ int64_t cell_idx(const CellEnt& c, int p)
{
    const int x = p >> 6, y = (p >> 3) & 7, z = p & 7;
    if (x < c.lo0 || x >= c.hi0 || y < c.lo1 || y >= c.hi1 || z < c.lo2 || z >= c.hi2) return -1;
    return c.base + ((int64_t)(c.b0 + x) * c.e + (c.b1 + y)) * c.e + (c.b2 + z);
}

// index of the box point at cell-local (x, y, z), or -1 outside the box
__device__ __forceinline__ int64_t cell_idx3(const CellEnt& c, int x, int y, int z)
{
    if (x < c.lo0 || x >= c.hi0 || y < c.lo1 || y >= c.hi1 || z < c.lo2 || z >= c.hi2) return -1;
    return c.base + ((int64_t)(c.b0 + x) * c.e + (c.b1 + y)) * c.e + (c.b2 + z);
}

constexpr int RJB = 2;  // point tiles per load batch in k_cell_residual and k_cell_density
constexpr int RJP = 1;  // 32-point tiles per load batch in k_cell_residual with paired loads

// Per cell: G = A^T A and P = A^T B (A = psi, B = H psi, 512 x k) on MFMA.
// Four waves split the 512 points (two x-planes each), then combine in LDS
// in a fixed order; partials go to part[slot + a*k + b] (G) and
// part[nslot + slot + a*k + b] (P).  MFMA step s of wave w covers the four
// points x = 2w + s/16, y = (s/2)%8, z = 4(s%2) + lane/16.
// one 16 x 16 tile pair (ti, tj) of a cell's S / H partials, summed over
// this wave's 128 points; SAME: ti == tj, so the B operand of S is A's
template <int PREC, bool SAME>
__device__ __forceinline__ void cell_gram_tile(const double* psi, const double* hpsi, const CellEnt& A,
    const CellEnt& B, int wid, int r, double (&accG)[4], double (&accP)[4])
{
    if (PREC == MGX_PREC_F64)
    {
        d4c g = { 0., 0., 0., 0. }, h = { 0., 0., 0., 0. };
#pragma unroll 4
        for (int s = 0; s < 32; s++)
        {
            const int x = 2 * wid + (s >> 4), y = (s >> 1) & 7, z = 4 * (s & 1) + r;
            const int64_t ia = cell_idx3(A, x, y, z);
            const int64_t ib = SAME ? ia : cell_idx3(B, x, y, z);
            const double xa = ia >= 0 ? psi[ia] : 0.;
            const double xb = SAME ? xa : (ib >= 0 ? psi[ib] : 0.);
            const double hb = ib >= 0 ? hpsi[ib] : 0.;
            g = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, xb, g, 0, 0, 0);
            h = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, hb, h, 0, 0, 0);
        }
        for (int v = 0; v < 4; v++)
        {
            accG[v] = g[v];
            accP[v] = h[v];
        }
    }
    else
    {
        // hi/lo split: hi*hi + (hi*lo + lo*hi) on FP32 MFMA; FP32 sums
        // over 8 steps (32 points), FP64 across them
        for (int v = 0; v < 4; v++)
            accG[v] = accP[v] = 0.;
        for (int s0 = 0; s0 < 32; s0 += 8)
        {
            f4c g1 = { 0.f, 0.f, 0.f, 0.f }, g2 = g1, h1 = g1, h2 = g1;
#pragma unroll 4
            for (int s = s0; s < s0 + 8; s++)
            {
                const int x = 2 * wid + (s >> 4), y = (s >> 1) & 7, z = 4 * (s & 1) + r;
                const int64_t ia = cell_idx3(A, x, y, z);
                const int64_t ib = SAME ? ia : cell_idx3(B, x, y, z);
                const double xa = ia >= 0 ? psi[ia] : 0.;
                const double xb = SAME ? xa : (ib >= 0 ? psi[ib] : 0.);
                const double hb = ib >= 0 ? hpsi[ib] : 0.;
                const float ah = (float)xa, al = (float)(xa - (double)ah);
                const float bh = SAME ? ah : (float)xb, bl = SAME ? al : (float)(xb - (double)bh);
                const float hh = (float)hb, hl = (float)(hb - (double)hh);
                g1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bh, g1, 0, 0, 0);
                g2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bl, g2, 0, 0, 0);
                g2 = __builtin_amdgcn_mfma_f32_16x16x4f32(al, bh, g2, 0, 0, 0);
                h1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, hh, h1, 0, 0, 0);
                h2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, hl, h2, 0, 0, 0);
                h2 = __builtin_amdgcn_mfma_f32_16x16x4f32(al, hh, h2, 0, 0, 0);
            }
            for (int v = 0; v < 4; v++)
            {
                accG[v] += (double)g1[v] + (double)g2[v];
                accP[v] += (double)h1[v] + (double)h2[v];
            }
        }
    }
}

// The same tile with each lane loading two adjacent z-points (16 bytes):
// lanes r = 0..3 of a box cover one 8-point cell row, and the two values feed
// two MFMA steps.  Valid when every box origin and edge is even, so the
// covered z-range of a cell row starts and ends at even positions and the
// pair is aligned.  The split keeps FP32 partial sums over 32 points.
// G > 1 packs G row groups into the 16 box columns (16/G boxes, k <= 16/G):
// lane column c = g (16/G) + a loads box a's row of group g, so one step
// covers G rows; the product's diagonal blocks (same g) are the partials.
// The cell's 64 rows of 8 points are summed in two fixed halves (rows 0-31,
// 32-63: NS steps of G rows each); one wave per cell computes both halves,
// two waves one each, and the halves are added in the same order either way,
// so the sums do not depend on the number of waves.
template <int PREC, bool SAME, int G>
__device__ __forceinline__ void cell_gram_tile2(const double* psi, const double* hpsi, const CellEnt& A,
    const CellEnt& B, int half, int r, int grp, double (&accG)[4], double (&accP)[4])
{
    constexpr int NS = 32 / G;
    const int wid    = half;
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 z2 = { 0., 0. };
    auto ld2 = [](const double* p, int64_t i) -> d2 { return *reinterpret_cast<const d2*>(p + i); };
    if (PREC == MGX_PREC_F64)
    {
        d4c g = { 0., 0., 0., 0. }, h = { 0., 0., 0., 0. };
#pragma unroll 4
        for (int s = 0; s < NS; s++)
        {
            const int row = (wid * NS + s) * G + grp;
            const int x = row >> 3, y = row & 7, z = 2 * r;
            const int64_t ia = cell_idx3(A, x, y, z);
            const int64_t ib = SAME ? ia : cell_idx3(B, x, y, z);
            const d2 xa = ia >= 0 ? ld2(psi, ia) : z2;
            const d2 xb = SAME ? xa : (ib >= 0 ? ld2(psi, ib) : z2);
            const d2 hb = ib >= 0 ? ld2(hpsi, ib) : z2;
            g = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[0], xb[0], g, 0, 0, 0);
            h = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[0], hb[0], h, 0, 0, 0);
            g = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[1], xb[1], g, 0, 0, 0);
            h = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[1], hb[1], h, 0, 0, 0);
        }
        for (int v = 0; v < 4; v++)
        {
            accG[v] = g[v];
            accP[v] = h[v];
        }
    }
    else
    {
        for (int v = 0; v < 4; v++)
            accG[v] = accP[v] = 0.;
        for (int s0 = 0; s0 < NS; s0 += 4)
        {
            f4c g1 = { 0.f, 0.f, 0.f, 0.f }, g2 = g1, h1 = g1, h2 = g1;
#pragma unroll 2
            for (int s = s0; s < s0 + 4; s++)
            {
                const int row = (wid * NS + s) * G + grp;
                const int x = row >> 3, y = row & 7, z = 2 * r;
                const int64_t ia = cell_idx3(A, x, y, z);
                const int64_t ib = SAME ? ia : cell_idx3(B, x, y, z);
                const d2 xa2 = ia >= 0 ? ld2(psi, ia) : z2;
                const d2 xb2 = SAME ? xa2 : (ib >= 0 ? ld2(psi, ib) : z2);
                const d2 hb2 = ib >= 0 ? ld2(hpsi, ib) : z2;
#pragma unroll
                for (int q = 0; q < 2; q++)
                {
                    const double xa = xa2[q], xb = xb2[q], hb = hb2[q];
                    const float ah = (float)xa, al = (float)(xa - (double)ah);
                    const float bh = SAME ? ah : (float)xb, bl = SAME ? al : (float)(xb - (double)bh);
                    const float hh = (float)hb, hl = (float)(hb - (double)hh);
                    g1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bh, g1, 0, 0, 0);
                    g2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bl, g2, 0, 0, 0);
                    g2 = __builtin_amdgcn_mfma_f32_16x16x4f32(al, bh, g2, 0, 0, 0);
                    h1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, hh, h1, 0, 0, 0);
                    h2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, hl, h2, 0, 0, 0);
                    h2 = __builtin_amdgcn_mfma_f32_16x16x4f32(al, hh, h2, 0, 0, 0);
                }
            }
            for (int v = 0; v < 4; v++)
            {
                accG[v] += (double)g1[v] + (double)g2[v];
                accP[v] += (double)h1[v] + (double)h2[v];
            }
        }
    }
}

// A cell with k <= 16/G boxes as one packed tile (G row groups per step):
// the partial (a, b) is the sum over the groups g of the product's entry
// (i, j) = (g kp + a, g kp + b), kp = 16/G, read from the lanes that hold it
// (FP64 MFMA: lane j + 16 (i % 4), value i / 4; FP32: lane j + 16 (i / 4),
// value i % 4), over the waves and groups in a fixed order.
template <int PREC, int NW, int G>
__device__ __forceinline__ void cell_gram_packed(const double* psi, const double* hpsi, const CellEnt* ce, int k,
    int wid, int lane, double (*red)[64][4], double* part, int64_t nslot)
{
    constexpr int kp = 16 / G;
    constexpr int HW = 2 / NW;  // halves per wave
    const int c = lane & 15, r = lane >> 4;
    const int a = c % kp, g = c / kp;
    CellEnt none{};
    const CellEnt A = a < k ? ce[a] : none;
    // red[pass * 2 + half]: each half's partials go to LDS as soon as formed
    for (int q = 0; q < HW; q++)
    {
        double accG[4], accP[4];
        const int h = wid * HW + q;
        cell_gram_tile2<PREC, true, G>(psi, hpsi, A, A, h, r, g, accG, accP);
        for (int v = 0; v < 4; v++)
        {
            red[h][lane][v]     = accG[v];
            red[2 + h][lane][v] = accP[v];
        }
    }
    __syncthreads();
    for (int pass = 0; pass < 2; pass++)
        for (int o = threadIdx.x; o < k * k; o += 64 * NW)
        {
            const int aa = o / k, bb = o - (o / k) * k;
            double sum = 0.;
            for (int w = 0; w < 2; w++)
                for (int gg = 0; gg < G; gg++)
                {
                    const int i = gg * kp + aa, j = gg * kp + bb;
                    const int ln = j + 16 * (PREC == MGX_PREC_F64 ? (i & 3) : (i >> 2));
                    sum += red[2 * pass + w][ln][PREC == MGX_PREC_F64 ? (i >> 2) : (i & 3)];
                }
            part[pass * nslot + o] = sum;
        }
}

// S / H partials per 8^3 cell over all box pairs covering it (k <= 32):
// G_ab = sum_p psi_a psi_b, P_ab = sum_p psi_a hpsi_b, as 16 x 16 box tiles
// on MFMA with K = points (4 per step); the NW waves split the cell's rows,
// and their sums are added in a fixed order.  With paired loads, cells of
// k <= 8 boxes pack 2 or 4 row groups into the tile's 16 columns, so that
// the loads of a step fill the wave (a cell is covered by 5-6 boxes on
// average when 32^3 boxes cover the grid 4 times).
template <int PREC, bool PAIR, int NW>
__global__ void __launch_bounds__(64 * NW, 5) k_cell_gram(const double* psi, const double* hpsi, const int* cell_ptr,
    const CellEnt* cce, const int64_t* cell_slot, int64_t nslot, double* part, int ncell)
{
    static_assert(PAIR || NW == 4, "the unpaired tiles take four waves");
    static_assert(!PAIR || NW == 1 || NW == 2, "the paired tiles take one or two waves");
    // paired tiles: one slot per half of the cell's rows; unpaired: per wave
    constexpr int NR = PAIR ? 2 : NW;
    __shared__ CellEnt ce[kCellK];
    __shared__ double red[2 * NR][64][4];  // [pass * NR + part]
    // consecutive cells (z-neighbours share cache lines) on one XCD
    const int64_t cl = xcd_block(blockIdx.x, gridDim.x);
    if (cl >= ncell) return;
    const int cell = (int)cl;
    const int e0 = cell_ptr[cell], k = cell_ptr[cell + 1] - e0;
    load_cell(cce + e0, k, ce);
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c = lane & 15, r = lane >> 4;
    const int nt = (k + 15) / 16;
    const int64_t slot = cell_slot[cell];
    CellEnt none{};
    if constexpr (PAIR)
    {
        if (k <= 8)
        {
            // k <= 4: four row groups per step, k <= 8: two
            if (k <= 4)
                cell_gram_packed<PREC, NW, 4>(psi, hpsi, ce, k, wid, lane, red, part + slot, nslot);
            else
                cell_gram_packed<PREC, NW, 2>(psi, hpsi, ce, k, wid, lane, red, part + slot, nslot);
            return;
        }
    }
    for (int ti = 0; ti < nt; ti++)
        for (int tj = 0; tj < nt; tj++)
        {
            const int a = ti * 16 + c, b = tj * 16 + c;
            const CellEnt A = a < k ? ce[a] : none;
            const CellEnt B = b < k ? ce[b] : none;
            constexpr int HW = PAIR ? (NW == 1 ? 2 : 1) : 1;  // row halves per wave (paired), else one part
            for (int q = 0; q < HW; q++)
            {
                double accG[4], accP[4];
                const int part_i = wid * HW + q;
                if constexpr (PAIR)
                {
                    if (ti == tj)
                        cell_gram_tile2<PREC, true, 1>(psi, hpsi, A, B, part_i, r, 0, accG, accP);
                    else
                        cell_gram_tile2<PREC, false, 1>(psi, hpsi, A, B, part_i, r, 0, accG, accP);
                }
                else if (ti == tj)
                    cell_gram_tile<PREC, true>(psi, hpsi, A, B, wid, r, accG, accP);
                else
                    cell_gram_tile<PREC, false>(psi, hpsi, A, B, wid, r, accG, accP);
                for (int v = 0; v < 4; v++)
                {
                    red[part_i][lane][v]      = accG[v];
                    red[NR + part_i][lane][v] = accP[v];
                }
            }
            __syncthreads();
            for (int pass = 0; pass < 2; pass++)
            {
                if (wid == 0)
                {
                    for (int v = 0; v < 4; v++)
                    {
                        double sum = red[pass * NR][lane][v];
                        for (int w = 1; w < NR; w++)
                            sum += red[pass * NR + w][lane][v];
                        const int row = ti * 16 + ((PREC == MGX_PREC_F64) ? r + 4 * v : 4 * r + v);
                        const int col = tj * 16 + c;
                        if (row < k && col < k) part[pass * nslot + slot + row * k + col] = sum;
                    }
                }
                __syncthreads();
            }
        }
}

// Stencil reach of H psi_j outside box j, organised by face: block (j, f)
// with f = d*2 + side covers the two slabs (planes at distance 1 and 2
// beyond the face; slab index s = f*2 + layer-1).  The reach values hv(X) of
// both planes are computed once into LDS; then every interacting box i
// (one wave each) contributes
//   out[12*q + s] = sum_{X in slab s, X in box i} psi_i(X) hv(X)
// for its entry q in j's interaction list, reading box i's two planes
// together (adjacent values for z faces).
// Blocks run the boxes in Morton order of their centres (order), mapped
// XCD-aware: the faces of nearby boxes, which read rows of the same
// neighbours, run on one XCD and share its L2.
__global__ void __launch_bounds__(256) k_slab_skin(const double* psi, const int64_t* off, const int* edge,
    const int* ilist_ptr, const int* ilist, Coef4 c4, int use_lds, double* out, const int* order, int nbox)
{
    extern __shared__ double hv[];  // 2 planes of ej^2
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    if (L >= 6 * (int64_t)nbox) return;
    const int j = order[L / 6], f = (int)(L - (L / 6) * 6);
    const int d = f >> 1, side = f & 1;
    const int ej = edge[j];
    const double* pj = psi + off[j];
    const double c1 = d == 0 ? c4.c1x : (d == 1 ? c4.c1y : c4.c1z);
    const double c2 = d == 0 ? c4.c2x : (d == 1 ? c4.c2y : c4.c2z);
    const int u = (d + 1) % 3, w = (d + 2) % 3;
    // plane positions along d (j-local) of layers 1 and 2
    const int Xd1 = side == 0 ? -1 : ej, Xd2 = side == 0 ? -2 : ej + 1;
    // the two box values next to the face: f0 at the face, f1 one further in
    auto reach = [&](int a, int b, double* h1, double* h2) {
        int q[3];
        q[u]        = a;
        q[w]        = b;
        q[d]        = side == 0 ? 0 : ej - 1;
        const double f0 = pj[((int64_t)q[0] * ej + q[1]) * ej + q[2]];
        double f1       = 0.;
        if (ej > 1)
        {
            q[d] = side == 0 ? 1 : ej - 2;
            f1   = pj[((int64_t)q[0] * ej + q[1]) * ej + q[2]];
        }
        *h1 = c1 * f0 + c2 * f1;
        *h2 = c2 * f0;
    };
    const int e2 = ej * ej;
    for (int t = threadIdx.x; use_lds && t < e2; t += blockDim.x)
    {
        const int a = t / ej, b = t - (t / ej) * ej;
        reach(a, b, hv + t, hv + e2 + t);
    }
    __syncthreads();
    // one wave per interacting box: no block barriers in the loop
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int s1 = 2 * f, s2 = 2 * f + 1;
    for (int qq = ilist_ptr[j] + wid; qq < ilist_ptr[j + 1]; qq += nwv)
    {
        const int* E = ilist + 4 * qq;
        const int i = E[0], ei = edge[i];
        // only <psi_i, H psi_j> with i < j is used (k_pair_reduce)
        if (i >= j) continue;
        const int sh[3] = { E[1], E[2], E[3] };
        const int xd1 = Xd1 + sh[d], xd2 = Xd2 + sh[d];
        const int lo_u = max(0, -sh[u]), hi_u = min(ej, ei - sh[u]);
        const int lo_w = max(0, -sh[w]), hi_w = min(ej, ei - sh[w]);
        const int nu = hi_u - lo_u, nww = hi_w - lo_w;
        // box i must also lie across the face (u and w ranges not empty:
        // boxes that are apart along u or w are in the list when they come
        // within 2 points along every axis, and two negative extents would
        // make a positive count)
        const bool across = nu > 0 && nww > 0;
        const bool in1 = across && xd1 >= 0 && xd1 < ei, in2 = across && xd2 >= 0 && xd2 < ei;
        double acc1 = 0., acc2 = 0.;
        if (in1 || in2)
        {
            const double* pi = psi + off[i];
            for (int t = lane; t < nu * nww; t += 64)
            {
                const int a = lo_u + t / nww, b = lo_w + (t - (t / nww) * nww);
                int x[3];
                x[u] = a + sh[u];
                x[w] = b + sh[w];
                double h1, h2;
                if (use_lds)
                {
                    h1 = hv[a * ej + b];
                    h2 = hv[e2 + a * ej + b];
                }
                else
                    reach(a, b, &h1, &h2);
                if (in1)
                {
                    x[d] = xd1;
                    acc1 += pi[((int64_t)x[0] * ei + x[1]) * ei + x[2]] * h1;
                }
                if (in2)
                {
                    x[d] = xd2;
                    acc2 += pi[((int64_t)x[0] * ei + x[1]) * ei + x[2]] * h2;
                }
            }
            for (int o = 32; o > 0; o >>= 1)
            {
                acc1 += __shfl_xor(acc1, o, 64);
                acc2 += __shfl_xor(acc2, o, 64);
            }
        }
        if (lane == 0)
        {
            out[12 * (int64_t)qq + s1] = acc1;
            out[12 * (int64_t)qq + s2] = acc2;
        }
    }
}

// S_ij, H_ij from the cell partials (slots in a fixed order) plus the skin
__global__ void k_pair_reduce(const double* part, int64_t nslot, const int* pslot_ptr, const int64_t* pslot,
    const double* skin, const int* pair_ient, const int* pairs, int64_t npairs, int n, double vel, double* S, double* H,
    double* Sv, double* Hv)
{
    // eight lanes per pair: lane k sums the slots k, k + 8, ..., then the
    // eight partial sums meet in a fixed butterfly
    const int sub = threadIdx.x & 7;
    for (int64_t t = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 3; t < npairs;
         t += ((int64_t)gridDim.x * blockDim.x) >> 3)
    {
        double g = 0., h1 = 0.;
        for (int q = pslot_ptr[t] + sub; q < pslot_ptr[t + 1]; q += 8)
        {
            const int64_t ab = pslot[2 * q];
            g += part[ab];
            h1 += part[nslot + ab];
        }
        for (int o = 4; o > 0; o >>= 1)
        {
            g += __shfl_xor(g, o, 64);
            h1 += __shfl_xor(h1, o, 64);
        }
        if (sub) continue;
        const int i = pairs[11 * t], j = pairs[11 * t + 1];
        if (i != j)
        {
            // h1 = <psi_i, H psi_j> (i < j): the in-box part plus the skin of
            // j seen by i (entry of i in j's list).  The stencil is symmetric,
            // so <psi_j, H psi_i> is the same number up to rounding; H takes
            // h1 on both sides and is exactly symmetric.
            const int q1 = pair_ient[2 * t];
            for (int sl = 0; sl < 12; sl++)
                h1 += skin[12 * (int64_t)q1 + sl];
        }
        const double sv = g * vel, hv = h1 * vel;
        S[(int64_t)j * n + i] = sv;
        S[(int64_t)i * n + j] = sv;
        H[(int64_t)j * n + i] = hv;
        H[(int64_t)i * n + j] = hv;
        if (Sv)
        {
            Sv[t] = sv;
            Hv[t] = hv;
        }
    }
}

// R_b(p) = sum_a psi_a(p) theta(a,b) - hpsi_b(p) for every box b covering
// point p of the cell, as a 512 x k by k x k product on FP64 MFMA:
// A = psi gathered per (point, box a), B = theta sub-block from LDS.
// where the residual goes: OM 0 writes R (FP64); OM 1 / 2 write instead the
// V-cycle's zero start in FP32 / FP64, f = (T)R and x = 0 + (T)(jf f), the
// values k_first_sweep forms from R (one pass over the points less)
template <int OM>
__device__ __forceinline__ void res_put(double* res, void* fo, void* xo, double jf, int64_t i, double rv)
{
    if constexpr (OM == 0)
        res[i] = rv;
    else if constexpr (OM == 1)
    {
        const float fv = (float)rv;
        float xv       = 0.f;
        xv += static_cast<float>(jf * static_cast<double>(fv));
        static_cast<float*>(fo)[i] = fv;
        static_cast<float*>(xo)[i] = xv;
    }
    else
    {
        double xv = 0.;
        xv += jf * rv;
        static_cast<double*>(fo)[i] = rv;
        static_cast<double*>(xo)[i] = xv;
    }
}

// two adjacent points (i even): 16-byte stores of R, 8-byte (FP32) or
// 16-byte (FP64) stores of the start
template <int OM>
__device__ __forceinline__ void res_put2(double* res, void* fo, void* xo, double jf, int64_t i, double r0, double r1)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef float f2 __attribute__((ext_vector_type(2)));
    if constexpr (OM == 0)
    {
        d2 v;
        v[0]                                = r0;
        v[1]                                = r1;
        *reinterpret_cast<d2*>(res + i)     = v;
    }
    else if constexpr (OM == 1)
    {
        f2 fv, xv;
        fv[0] = (float)r0;
        fv[1] = (float)r1;
        xv[0] = 0.f;
        xv[1] = 0.f;
        xv[0] += static_cast<float>(jf * static_cast<double>(fv[0]));
        xv[1] += static_cast<float>(jf * static_cast<double>(fv[1]));
        *reinterpret_cast<f2*>(static_cast<float*>(fo) + i) = fv;
        *reinterpret_cast<f2*>(static_cast<float*>(xo) + i) = xv;
    }
    else
    {
        d2 fv, xv;
        fv[0] = r0;
        fv[1] = r1;
        xv[0] = 0.;
        xv[1] = 0.;
        xv[0] += jf * r0;
        xv[1] += jf * r1;
        *reinterpret_cast<d2*>(static_cast<double*>(fo) + i) = fv;
        *reinterpret_cast<d2*>(static_cast<double*>(xo) + i) = xv;
    }
}

// The outputs never alias the inputs.  A block reads its cell's entries (one
// dependent load after the cell pointer); with paired loads and k <= 16 each
// lane then loads its theta operands itself, at the same time as the first
// psi / H psi batch (no LDS stage and no second barrier); otherwise theta goes
// through LDS.
template <bool PAIR, int OM>
__global__ void __launch_bounds__(256, 4) k_cell_residual(const double* __restrict__ psi,
    const double* __restrict__ hpsi, const double* __restrict__ theta, int n, const int* cell_ptr, const CellEnt* cce,
    double* __restrict__ res, double* __restrict__ esq, int ncell, void* __restrict__ fo, void* __restrict__ xo,
    double jf)
{
    __shared__ CellEnt ce[kCellK];
    __shared__ double th[kCellK * kCellK];
    __shared__ double wsq[4][kCellK];
    // consecutive cells (z-neighbours share cache lines) on one XCD
    const int64_t cl = xcd_block(blockIdx.x, gridDim.x);
    if (cl >= ncell) return;
    const int cell = (int)cl;
    const int e0 = cell_ptr[cell], k = cell_ptr[cell + 1] - e0;
    load_cell(cce + e0, k, ce);
    __syncthreads();
    if (!PAIR || k > 16)
    {
        for (int t = threadIdx.x; t < kCellK * kCellK; t += blockDim.x)
        {
            const int aa = t / kCellK, bb = t - (t / kCellK) * kCellK;
            th[t] = (aa < k && bb < k) ? theta[(int64_t)ce[bb].box * n + ce[aa].box] : 0.;
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c = lane & 15, r = lane >> 4;
    const int nct = (k + 15) / 16;
    const int ksteps = (k + 3) / 4;
    // this lane's sums of R^2 for boxes 16 t + r + 4 v (esq: per cell entry)
    double sq[2][4] = { { 0., 0., 0., 0. }, { 0., 0., 0., 0. } };
    if (PAIR && k <= 16)
    {
        // even box origins and edges: each lane takes two adjacent z-points
        // (16-byte loads and stores), which feed two MFMAs with the same
        // theta operand; 32-point tiles, wave w takes tiles w + 4 jb.  The
        // loads of tile jb + 1 are issued before the products of tile jb
        // (two register sets), and those of the first tile before theta's.
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 z2 = { 0., 0. };
        // psi of boxes 4 st + r and H psi of boxes r + 4 v at the lane's pair
        auto issue = [&](int jb, d2(&xa)[4], d2(&hv)[4]) {
            const int pp = (wid + 4 * jb) * 32 + 2 * c;
#pragma unroll
            for (int st = 0; st < 4; st++)
            {
                const int aa = 4 * st + r;
                xa[st]       = z2;
                if (aa < k)
                {
                    const int64_t ia = cell_idx(ce[aa], pp);
                    if (ia >= 0) xa[st] = *reinterpret_cast<const d2*>(psi + ia);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int bb = r + 4 * v;
                hv[v]        = z2;
                if (bb < k)
                {
                    const int64_t ib = cell_idx(ce[bb], pp);
                    if (ib >= 0) hv[v] = *reinterpret_cast<const d2*>(hpsi + ib);
                }
            }
        };
        d2 xa[2][4], hv[2][4];
        issue(0, xa[0], hv[0]);
        for (int t = threadIdx.x; t < kCellK * kCellK; t += blockDim.x)
        {
            const int aa = t / kCellK, bb = t - (t / kCellK) * kCellK;
            th[t] = (aa < k && bb < k) ? theta[(int64_t)ce[bb].box * n + ce[aa].box] : 0.;
        }
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < 4; jb++)
        {
            const int cur = jb & 1;
            if (jb + 1 < 4) issue(jb + 1, xa[cur ^ 1], hv[cur ^ 1]);
            const int pp = (wid + 4 * jb) * 32 + 2 * c;
            d4c acc0 = { 0., 0., 0., 0. }, acc1 = acc0;
#pragma unroll
            for (int st = 0; st < 4; st++)
                if (st < ksteps)
                {
                    const double t = th[(4 * st + r) * kCellK + c];
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(t, xa[cur][st][0], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(t, xa[cur][st][1], acc1, 0, 0, 0);
                }
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int bb = r + 4 * v;
                if (bb >= k) continue;
                const int64_t ib = cell_idx(ce[bb], pp);
                if (ib >= 0)
                {
                    d2 rv;
                    rv[0] = acc0[v] + (-1. * hv[cur][v][0]);
                    rv[1] = acc1[v] + (-1. * hv[cur][v][1]);
                    res_put2<OM>(res, fo, xo, jf, ib, rv[0], rv[1]);
                    sq[0][v] += rv[0] * rv[0];
                    sq[0][v] += rv[1] * rv[1];
                }
            }
        }
    }
    else if (k <= 16)
    {
        // one box tile: the wave's 8 point tiles in batches of RJB, every
        // load of a batch issued before its products (same sums, same order)
        for (int jb = 0; jb < 8; jb += RJB)
        {
            double xa[RJB][4], hv[RJB][4];
            int64_t ib[RJB][4];
#pragma unroll
            for (int u = 0; u < RJB; u++)
            {
                const int pp = (wid + 4 * (jb + u)) * 16 + c;
#pragma unroll
                for (int st = 0; st < 4; st++)
                {
                    const int aa = 4 * st + r;
                    xa[u][st]    = 0.;
                    if (aa < k)
                    {
                        const int64_t ia = cell_idx(ce[aa], pp);
                        if (ia >= 0) xa[u][st] = psi[ia];
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; v++)
                {
                    const int bb = r + 4 * v;
                    ib[u][v]     = bb < k ? cell_idx(ce[bb], pp) : -1;
                    hv[u][v]     = ib[u][v] >= 0 ? hpsi[ib[u][v]] : 0.;
                }
            }
#pragma unroll
            for (int u = 0; u < RJB; u++)
            {
                d4c acc = { 0., 0., 0., 0. };
#pragma unroll
                for (int st = 0; st < 4; st++)
                    if (st < ksteps)
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(th[(4 * st + r) * kCellK + c], xa[u][st], acc, 0,
                            0, 0);
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (ib[u][v] >= 0)
                    {
                        const double rv = acc[v] + (-1. * hv[u][v]);
                        res_put<OM>(res, fo, xo, jf, ib[u][v], rv);
                        sq[0][v] += rv * rv;
                    }
            }
        }
    }
    for (int job = wid; k > 16 && job < 32 * nct; job += 4)
    {
        const int p0 = (job & 31) * 16, b0 = (job >> 5) * 16;
        d4c acc = { 0., 0., 0., 0. };
        for (int st = 0; st < ksteps; st++)
        {
            const int aa  = 4 * st + r;
            const int pp  = p0 + c;
            double xa     = 0.;
            if (aa < k)
            {
                const int64_t ia = cell_idx(ce[aa], pp);
                if (ia >= 0) xa = psi[ia];
            }
            const double tb = th[aa * kCellK + b0 + c];
            // D[m = box b0+m][n = point p0+n] = sum_a theta(a, b) psi_a(point):
            // A = theta^T (lane: m = c, k = r), B = psi (k = r, n = c)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(tb, xa, acc, 0, 0, 0);
        }
        // lane (c, r) holds boxes b0 + r + 4v at point p0 + c: 16 lanes write
        // 16 consecutive points of one box
        const int pp = p0 + c;
        for (int v = 0; v < 4; v++)
        {
            const int bb = b0 + r + 4 * v;
            if (bb >= k) continue;
            const int64_t ib = cell_idx(ce[bb], pp);
            if (ib >= 0)
            {
                const double rv = acc[v] + (-1. * hpsi[ib]);
                res_put<OM>(res, fo, xo, jf, ib, rv);
                sq[job >> 5][v] += rv * rv;
            }
        }
    }
    if (!esq) return;
    // sums over the 16 points of a lane group, then over the waves in order
    for (int t = 0; t < 2; t++)
        for (int v = 0; v < 4; v++)
        {
            double x = sq[t][v];
            for (int o = 1; o < 16; o <<= 1)
                x += __shfl_xor(x, o, 64);
            const int bb = 16 * t + r + 4 * v;
            if (c == 0 && bb < kCellK) wsq[wid][bb] = x;
        }
    __syncthreads();
    for (int b = threadIdx.x; b < k; b += blockDim.x)
        esq[e0 + b] = ((wsq[0][b] + wsq[1][b]) + wsq[2][b]) + wsq[3][b];
}

// d_j = vel * sum of the box's cell-entry partials (entries listed per box
// in cell order: a fixed summation order)
__global__ void __launch_bounds__(64) k_box_from_entries(const double* esq, const int* bent_ptr, const int* bent,
    int nbox, double vel, double* d)
{
    const int j = blockIdx.x;
    if (j >= nbox) return;
    double acc = 0.;
    for (int q = bent_ptr[j] + threadIdx.x; q < bent_ptr[j + 1]; q += 64)
        acc += esq[bent[q]];
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, 64);
    if (threadIdx.x == 0) d[j] = acc * vel;
}

// Electron density on the cells (Rho::computeRhoSubdomainUsingBlas3,
// src/Rho.cc:360-449: product = Phi X, then rho += product o Phi):
//   rho(p) = sum_b psi_b(p) sum_a X_ab psi_a(p)
// over the k boxes covering the cell, the inner sum as the (k x k)(k x 512)
// product on FP64 MFMA (as in k_cell_residual), the outer one per lane and
// over the four lanes of a point.  Each point of the periodic grid lies in
// one cell, so rho is written once per point (points no box covers stay as
// the caller set them); csum[cell] = the cell's sum of rho, in a fixed order.
template <bool PAIR>
__global__ void __launch_bounds__(256) k_cell_density(const double* psi, const double* X, int n, const int* origin, int g0,
    int g1, int g2, const int* cell_ptr, const CellEnt* cce, double* rho, double* csum, int ncell)
{
    __shared__ CellEnt ce[kCellK];
    __shared__ double xs[kCellK * kCellK];
    __shared__ int jid[kCellK];
    __shared__ double wsum[4];
    const int64_t cl = xcd_block(blockIdx.x, gridDim.x);
    if (cl >= ncell) return;
    const int cell = (int)cl;
    const int e0 = cell_ptr[cell], k = cell_ptr[cell + 1] - e0;
    load_cell(cce + e0, k, ce);
    for (int t = threadIdx.x; t < k; t += blockDim.x)
        jid[t] = cce[e0 + t].box;
    __syncthreads();
    for (int t = threadIdx.x; t < kCellK * kCellK; t += blockDim.x)
    {
        const int aa = t / kCellK, bb = t - (t / kCellK) * kCellK;
        xs[t] = (aa < k && bb < k) ? X[(int64_t)jid[bb] * n + jid[aa]] : 0.;
    }
    __syncthreads();
    // global coordinates of the cell origin (from any covering box)
    int gc[3];
    {
        const int gd[3] = { g0, g1, g2 };
        const int bl[3] = { ce[0].b0, ce[0].b1, ce[0].b2 };
        for (int d = 0; d < 3; d++)
        {
            int v = (origin[3 * jid[0] + d] + bl[d]) % gd[d];
            gc[d] = v < 0 ? v + gd[d] : v;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c = lane & 15, r = lane >> 4;
    const int nct = (k + 15) / 16;
    const int ksteps = (k + 3) / 4;
    double wacc = 0.;
    if (PAIR && k <= 16)
    {
        // even box origins and edges: two adjacent z-points per lane (16-byte
        // loads and stores), two MFMAs per X operand; 32-point tiles, wave w
        // takes tiles w + 4 jb
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 z2 = { 0., 0. };
#pragma unroll 1
        for (int jb = 0; jb < 4; jb++)
        {
            const int pp = (wid + 4 * jb) * 32 + 2 * c;
            d2 xa[4];
#pragma unroll
            for (int st = 0; st < 4; st++)
            {
                const int aa = 4 * st + r;
                xa[st]       = z2;
                if (aa < k)
                {
                    const int64_t ia = cell_idx(ce[aa], pp);
                    if (ia >= 0) xa[st] = *reinterpret_cast<const d2*>(psi + ia);
                }
            }
            d4c acc0 = { 0., 0., 0., 0. }, acc1 = acc0;
#pragma unroll
            for (int st = 0; st < 4; st++)
                if (st < ksteps)
                {
                    const double t = xs[(4 * st + r) * kCellK + c];
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(t, xa[st][0], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(t, xa[st][1], acc1, 0, 0, 0);
                }
            double rs0 = 0., rs1 = 0.;
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (r + 4 * v < k)
                {
                    rs0 += acc0[v] * xa[v][0];
                    rs1 += acc1[v] * xa[v][1];
                }
            rs0 += __shfl_xor(rs0, 16, 64);
            rs1 += __shfl_xor(rs1, 16, 64);
            rs0 += __shfl_xor(rs0, 32, 64);
            rs1 += __shfl_xor(rs1, 32, 64);
            if (r == 0)
            {
                // the pair stays in the cell's z-row (cells are 8-aligned)
                int gx = gc[0] + (pp >> 6), gy = gc[1] + ((pp >> 3) & 7), gz = gc[2] + (pp & 7);
                if (gx >= g0) gx -= g0;
                if (gy >= g1) gy -= g1;
                if (gz >= g2) gz -= g2;
                d2 o;
                o[0] = rs0;
                o[1] = rs1;
                *reinterpret_cast<d2*>(rho + ((int64_t)gx * g1 + gy) * g2 + gz) = o;
                wacc += rs0;
                wacc += rs1;
            }
        }
    }
    // up to 16 boxes (one box tile): point tiles in batches of RJB with all
    // loads of a batch issued first (same sums in the same order)
    for (int jb = 0; !PAIR && k <= 16 && jb < 8; jb += RJB)
    {
        double xa[RJB][4];
#pragma unroll
        for (int u = 0; u < RJB; u++)
        {
            const int pp = (wid + 4 * (jb + u)) * 16 + c;
#pragma unroll
            for (int st = 0; st < 4; st++)
            {
                const int aa = 4 * st + r;
                xa[u][st]    = 0.;
                if (aa < k)
                {
                    const int64_t ia = cell_idx(ce[aa], pp);
                    if (ia >= 0) xa[u][st] = psi[ia];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < RJB; u++)
        {
            const int pp = (wid + 4 * (jb + u)) * 16 + c;
            d4c acc      = { 0., 0., 0., 0. };
#pragma unroll
            for (int st = 0; st < 4; st++)
                if (st < ksteps)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xs[(4 * st + r) * kCellK + c], xa[u][st], acc, 0, 0, 0);
            // the outer factor psi_b(p), b = r + 4v, is this lane's xa[u][v]
            double rs = 0.;
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (r + 4 * v < k) rs += acc[v] * xa[u][v];
            rs += __shfl_xor(rs, 16, 64);
            rs += __shfl_xor(rs, 32, 64);
            if (r == 0)
            {
                int gx = gc[0] + (pp >> 6), gy = gc[1] + ((pp >> 3) & 7), gz = gc[2] + (pp & 7);
                if (gx >= g0) gx -= g0;
                if (gy >= g1) gy -= g1;
                if (gz >= g2) gz -= g2;
                rho[((int64_t)gx * g1 + gy) * g2 + gz] = rs;
                wacc += rs;
            }
        }
    }
    for (int job = wid; k > 16 && job < 32; job += 4)
    {
        const int p0 = job * 16, pp = p0 + c;
        double rs = 0.;
        for (int bt = 0; bt < nct; bt++)
        {
            const int b0 = bt * 16;
            d4c acc = { 0., 0., 0., 0. };
            for (int st = 0; st < ksteps; st++)
            {
                const int aa = 4 * st + r;
                double xa    = 0.;
                if (aa < k)
                {
                    const int64_t ia = cell_idx(ce[aa], pp);
                    if (ia >= 0) xa = psi[ia];
                }
                const double tb = xs[aa * kCellK + b0 + c];
                // D[m = box b0+m][point p0+n] = sum_a X(a, b) psi_a(point)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(tb, xa, acc, 0, 0, 0);
            }
            // lane (c, r) holds boxes b0 + r + 4v at point p0 + c
            for (int v = 0; v < 4; v++)
            {
                const int bb = b0 + r + 4 * v;
                if (bb >= k) continue;
                const int64_t ib = cell_idx(ce[bb], pp);
                if (ib >= 0) rs += acc[v] * psi[ib];
            }
        }
        rs += __shfl_xor(rs, 16, 64);
        rs += __shfl_xor(rs, 32, 64);
        if (r == 0)
        {
            int gx = gc[0] + (pp >> 6), gy = gc[1] + ((pp >> 3) & 7), gz = gc[2] + (pp & 7);
            if (gx >= g0) gx -= g0;
            if (gy >= g1) gy -= g1;
            if (gz >= g2) gz -= g2;
            rho[((int64_t)gx * g1 + gy) * g2 + gz] = rs;
            wacc += rs;
        }
    }
    // lanes 0..15 hold the wave's point values: add them in lane order
    for (int o = 1; o < 16; o <<= 1)
        wacc += __shfl_xor(wacc, o, 64);
    if (lane == 0) wsum[wid] = wacc;
    __syncthreads();
    if (threadIdx.x == 0) csum[cell] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// out[0] = vel * sum_i v[i] in a fixed order (one block)
__global__ void __launch_bounds__(256) k_sum_scaled(const double* v, int n, double vel, double* out)
{
    __shared__ double lds[4];
    double acc[1] = { 0. };
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        acc[0] += v[i];
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) out[0] = acc[0] * vel;
}

__global__ void k_scale_d(double* u, int64_t n, double a)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        u[t] *= a;
}

__global__ void k_pair_scatter(const double* part, const int* pairs, int64_t npairs, int n, double vel, double* S,
    double* H, double* Sv, double* Hv)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < npairs; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int i = pairs[11 * t], j = pairs[11 * t + 1];
        const double s = part[3 * t] * vel;
        const double h = 0.5 * (part[3 * t + 1] + part[3 * t + 2]) * vel;
        S[(int64_t)j * n + i] = s;
        S[(int64_t)i * n + j] = s;
        H[(int64_t)j * n + i] = h;
        H[(int64_t)i * n + j] = h;
        if (Sv)
        {
            Sv[t] = s;
            Hv[t] = h;
        }
    }
}

// res_j = sum_k psi_k theta_kj - hpsi_j on box j (multiplyByMatrix + axpy),
// neighbours k in ascending order as the oracle accumulates them
__global__ void __launch_bounds__(256) k_box_residual(const double* psi, const double* hpsi, const double* theta,
    int n, const int* nbr_ptr, const int* nbr, BoxView v, double* res)
{
    int j, e, x, y, z, p;
    int64_t base;
    if (!tile_point(v, &j, &e, &x, &y, &z, &base, &p)) return;
    double acc = 0.;
    const int k0 = nbr_ptr[j], k1 = nbr_ptr[j + 1];
    for (int q = k0; q < k1; q++)
    {
        const int* N = nbr + 10 * q;
        if (x < N[1] || x >= N[4] || y < N[2] || y >= N[5] || z < N[3] || z >= N[6]) continue;
        const int k   = N[0];
        const int ek  = v.edge[k];
        const double t = theta[(int64_t)j * n + k];
        acc += psi[v.off[k] + ((int64_t)(x + N[7]) * ek + (y + N[8])) * ek + (z + N[9])] * t;
    }
    acc += -1. * hpsi[base + p];
    res[base + p] = acc;
}

// d_j = vel * sum r_j^2 : one block per box
__global__ void __launch_bounds__(256) k_box_sumsq(const double* r, const int64_t* off, int nbox, double vel,
    double* d)
{
    __shared__ double lds[4];
    const int j = blockIdx.x;
    if (j >= nbox) return;
    double acc[1] = { 0. };
    for (int64_t q = off[j] + threadIdx.x; q < off[j + 1]; q += blockDim.x)
        acc[0] += r[q] * r[q];
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) d[j] = acc[0] * vel;
}

// Anderson mixing inner products (AndersonMix::update, src/AndersonMix.cc,
// m = 1, with LocGridOrbitals::dotProductDiagonal): per box j,
//   dd_j = vel <f - f0, f - f0>,  df_j = vel <f - f0, f>
// with f = a*kr and f0 = a*kr0 formed in FP64 from the preconditioner's
// outputs of this and the previous step (type T, as the V-cycle returns them)
template <typename T>
__global__ void __launch_bounds__(256) k_box_and_dots(const T* kr, double a, const T* kr0, const int64_t* off,
    int nbox, double vel, double* d)
{
    __shared__ double lds[8];
    const int j = blockIdx.x;
    if (j >= nbox) return;
    double acc[2] = { 0., 0. };
    for (int64_t q = off[j] + threadIdx.x; q < off[j + 1]; q += blockDim.x)
    {
        const double f = a * (double)kr[q], df = f - a * (double)kr0[q];
        acc[0] += df * df;
        acc[1] += df * f;
    }
    block_sum<2>(acc, lds);
    if (threadIdx.x == 0)
    {
        d[j]        = acc[0] * vel;
        d[nbox + j] = acc[1] * vel;
    }
}

// out[k] = sum_j d[k*n + j] Sinv_jj, k = 0, 1 (single block)
__global__ void __launch_bounds__(256) k_diag_wsum2(const double* d, const double* Sinv, int n, double* out)
{
    __shared__ double lds[8];
    double acc[2] = { 0., 0. };
    for (int j = threadIdx.x; j < n; j += blockDim.x)
    {
        const double w = Sinv[(int64_t)j * n + j];
        acc[0] += d[j] * w;
        acc[1] += d[n + j] * w;
    }
    block_sum<2>(acc, lds);
    if (threadIdx.x == 0)
    {
        out[0] = acc[0];
        out[1] = acc[1];
        // AndersonMix's theta for one history vector: rhs / mat, out of
        // range values reset (above 0.5: 0, or -0.5 above 1), at least -3
        double th = acc[0] > 0. ? acc[1] / acc[0] : 0.;
        if (th > 0.5)
            th = th > 1. ? -0.5 : 0.;
        else if (th < -3.)
            th = -3.;
        out[2] = th;
    }
}

// Anderson m = 1 update (AndersonMix::update): x_bar = (1-th) x + th x0,
// f_bar = (1-th) f + th f0 with f = a*kr, f0 = a*kr0 (th from thp, kr0 not
// read when th = 0); kr0_out <- kr when non-null, and the new
// x = x_bar + beta f_bar written to xo (the buffer holding x0, which the
// caller then swaps with x: the old x becomes x0 without a copy), with the
// normalisation sums of the new x, d_j = vel |x_j|^2, in k_box_sumsq's
// order: one block per box
template <typename T>
__global__ void __launch_bounds__(256) k_box_and_update_sq(const double* x, double* xo, const T* kr, double a,
    const T* kr0, T* kr0_out, const int64_t* off, int nbox, const double* thp, double beta, double vel, double* d)
{
    __shared__ double lds[4];
    const int j = blockIdx.x;
    if (j >= nbox) return;
    const double th = thp ? *thp : 0.;
    double acc[1] = { 0. };
    for (int64_t q = off[j] + threadIdx.x; q < off[j + 1]; q += blockDim.x)
    {
        // without a history (th = 0) the history buffers are not read: they
        // may not hold values yet
        const T k       = kr[q];
        const double xv = x[q], fv = a * (double)k;
        const double xb = th != 0. ? (1. - th) * xv + th * xo[q] : xv;
        const double fb = th != 0. ? (1. - th) * fv + th * (a * (double)kr0[q]) : fv;
        if (kr0_out) kr0_out[q] = k;
        const double nv = xb + beta * fb;
        xo[q]           = nv;
        acc[0] += nv * nv;
    }
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) d[j] = acc[0] * vel;
}

// the Anderson combination of k_box_and_update_nrm for PT points per thread:
// nv = xb + beta fb, xb = (1 - th) x + th x0, fb = (1 - th) f + th f0 (f = a kr)
// with a history (HIST), nv = x + beta f without; acc += nv^2
template <typename T, int PT, bool HIST>
__device__ __forceinline__ void and_combine(const double* xj, const double* xoj, const T* kj, const T* k0j, T* k0o,
    double a, double th, double beta, int n, double* nv, double& acc)
{
#pragma unroll
    for (int c = 0; c < PT; c += 8)
    {
        // eight points in flight at a time (the scheduler may not hoist the
        // next group's loads: registers hold the PT values already formed)
#pragma unroll
        for (int i = c; i < c + 8; i++)
        {
            const int t = threadIdx.x + 1024 * i;
            nv[i]       = 0.;
            if (t < n)
            {
                const T k       = kj[t];
                const double xv = xj[t], fv = a * (double)k;
                double xb = xv, fb = fv;
                if constexpr (HIST)
                {
                    const double x0v = xoj[t], f0v = a * (double)k0j[t];
                    xb = (1. - th) * xv + th * x0v;
                    fb = (1. - th) * fv + th * f0v;
                }
                if (k0o) k0o[t] = k;
                nv[i] = xb + beta * fb;
                acc += nv[i] * nv[i];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// k_box_and_update_sq with the normalisation done in the same pass: one
// block of 1024 threads per box keeps the box's new values in registers
// (PT per thread, boxes of at most 1024 PT points), forms the box's norm
// d_j = vel |x_j|^2 and writes x_j / sqrt(d_j) (LocGridOrbitals::normalize,
// with no pending scale left for the next H psi)
template <typename T, int PT>
__global__ void __launch_bounds__(1024) k_box_and_update_nrm(const double* x, double* xo, const T* kr, double a,
    const T* kr0, T* kr0_out, const int64_t* off, int nbox, const double* thp, double beta, double vel, double* d)
{
    __shared__ double lds[17];
    const int j = blockIdx.x;
    if (j >= nbox) return;
    const double th = thp ? *thp : 0.;
    const int64_t q0 = off[j];
    const int n      = (int)(off[j + 1] - q0);
    const double* xj = x + q0;
    double* xoj      = xo + q0;
    const T* kj      = kr + q0;
    const T* k0j     = kr0 + q0;
    T* k0o           = kr0_out ? kr0_out + q0 : nullptr;
    double nv[PT];
    double acc[1] = { 0. };
    // without a history (th = 0) the history buffers are not read: they may
    // not hold values yet (th is the same for the whole block)
    if (th != 0.)
        and_combine<T, PT, true>(xj, xoj, kj, k0j, k0o, a, th, beta, n, nv, acc[0]);
    else
        and_combine<T, PT, false>(xj, xoj, kj, k0j, k0o, a, th, beta, n, nv, acc[0]);
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) lds[16] = acc[0] * vel;
    __syncthreads();
    const double dj = lds[16];
    const double fs = 1. / sqrt(dj);
#pragma unroll
    for (int i = 0; i < PT; i++)
    {
        const int t = threadIdx.x + 1024 * i;
        if (t < n) xoj[t] = nv[i] * fs;
    }
    if (threadIdx.x == 0) d[j] = dj;
}

// psi += a * kr with the normalisation sums of the new psi (k_box_sumsq's order)
template <typename T>
__global__ void __launch_bounds__(256) k_box_psi_step_sq(double* psi, const T* kr, double a, const int64_t* off,
    int nbox, double vel, double* d)
{
    __shared__ double lds[4];
    const int j = blockIdx.x;
    if (j >= nbox) return;
    double acc[1] = { 0. };
    for (int64_t q = off[j] + threadIdx.x; q < off[j + 1]; q += blockDim.x)
    {
        double v = psi[q];
        v += a * (double)kr[q];
        psi[q] = v;
        acc[0] += v * v;
    }
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) d[j] = acc[0] * vel;
}


// 0.5*sqrt(sum_j d_j Sinv_jj): single block
__global__ void __launch_bounds__(256) k_norm(const double* d, const double* Sinv, int n, double* out)
{
    __shared__ double lds[4];
    double acc[1] = { 0. };
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        acc[0] += d[j] * Sinv[(int64_t)j * n + j];
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) out[0] = 0.5 * sqrt(acc[0]);
}

// LocGridOrbitals::normalize, src/LocGridOrbitals.cc:2230-2290:
// psi_j *= 1/sqrt(vel*|psi_j|^2)
__global__ void k_box_scale(double* psi, const int64_t* off, const int* tile, int ntiles, const int* edge,
    const double* d)
{
    const int64_t t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= ntiles) return;
    const int j = tile[2 * t];
    const int p = tile[2 * t + 1] + threadIdx.x;
    const int e = edge[j];
    if (p >= e * e * e) return;
    psi[off[j] + p] *= 1. / sqrt(d[j]);
}

}  // namespace

// ----------------------------------------------------------- internals ---
int box_tiles_launch_check(const mgx_boxset* bs, int level)
{
    return (level >= 0 && level < bs->levels) ? MGX_OK : MGX_EINVAL;
}

// blocks of the per-box marching kernels: one per box, padded to a multiple of 8
inline unsigned march_grid_dim(const mgx_boxset* bs) { return (unsigned)((bs->nbox + 7) / 8 * 8); }

template <typename T, int TYPE, int MODE, int NT, int P>
void march_launch(mgx_ctx* ctx, const mgx_boxset* bs, const MarchArgs& a, const T* in, const T* f, const double* v,
    T* out, T* w, const Coefs& c)
{
    // two plane buffers unless that costs occupancy (8 blocks of this size
    // per CU must fit the 160 KB of LDS); loads two planes ahead when the
    // extra registers are few
    const size_t plane = (size_t)a.lds_w * a.lds_w * sizeof(T);
    constexpr int PD   = (P * sizeof(T) <= 16) ? 2 : 1;
    if (2 * plane * 8 <= 160 * 1024)
        hipLaunchKernelGGL((k_box_march<T, TYPE, MODE, NT, P, 2, PD>), dim3(march_grid_dim(bs)), dim3(NT), 2 * plane,
            ctx->stream, in, f, v, out, w, a, c.c4, c.c2);
    else
        hipLaunchKernelGGL((k_box_march<T, TYPE, MODE, NT, P, 1, PD>), dim3(march_grid_dim(bs)), dim3(NT), plane, ctx->stream,
            in, f, v, out, w, a, c.c4, c.c2);
}

// points per lane of k_box_march4: 16 bytes of FP32 / FP64, 4 FP16 values
// (8 bytes: 16-byte groups leave half the lanes of a 32^3 box idle)
template <typename T>
constexpr int march_v()
{
    return sizeof(T) == 2 ? 4 : 16 / (int)sizeof(T);
}

template <typename T, int TYPE, int MODE, int NT, int P, int XT, typename TO>
void march4_launch_d(mgx_ctx* ctx, const mgx_boxset* bs, const MarchArgs& a, size_t plane, const T* in, const T* f,
    const double* v, TO* out, T* w, const Coefs& c)
{
    constexpr int V  = march_v<T>();
    constexpr int ZR = V > 2 ? V : 2;
    // the fused restriction's two planes of w, or the fused prolongation's
    // three coarse planes, behind the stencil planes
    const int em    = a.lds_w - 2 * ZR;
    const size_t rb = a.rc ? 2 * (size_t)em * em * sizeof(T) : a.pc ? 3 * (size_t)a.cpl * sizeof(T) : 0;
    // two plane buffers when the blocks that fill a CU (32 waves) still fit
    // its 160 KB of LDS
    if (2 * plane * (2048 / NT) <= 160 * 1024)
        hipLaunchKernelGGL((k_box_march4<T, TYPE, MODE, NT, P, 2, V, XT, TO>), dim3(march_grid_dim(bs)), dim3(NT),
            2 * plane + rb, ctx->stream, in, f, v, out, w, a, c.c4, c.c2);
    else
        hipLaunchKernelGGL((k_box_march4<T, TYPE, MODE, NT, P, 1, V, XT, TO>), dim3(march_grid_dim(bs)), dim3(NT),
            plane + rb, ctx->stream, in, f, v, out, w, a, c.c4, c.c2);
}

// the Anderson dots (a.dout set) and the fused restriction (a.rc set, output
// type T) exist for the Jacobi sweep only, each in a kernel of its own
template <typename T, int TYPE, int MODE, int NT, int P, typename TO>
void march4_launch(mgx_ctx* ctx, const mgx_boxset* bs, const MarchArgs& a, size_t plane, const T* in, const T* f,
    const double* v, TO* out, T* w, const Coefs& c)
{
    constexpr int J = MODE == MARCH_JACOBI ? 1 : 0;
    if (J && a.dout)
    {
        march4_launch_d<T, TYPE, MODE, NT, P, J, TO>(ctx, bs, a, plane, in, f, v, out, w, c);
        return;
    }
    if constexpr (J && std::is_same<T, TO>::value)
    {
        if (a.rc)
        {
            march4_launch_d<T, TYPE, MODE, NT, P, 2, TO>(ctx, bs, a, plane, in, f, v, out, w, c);
            return;
        }
        if (a.pc)
        {
            march4_launch_d<T, TYPE, MODE, NT, P, 3, TO>(ctx, bs, a, plane, in, f, v, out, w, c);
            return;
        }
    }
    march4_launch_d<T, TYPE, MODE, NT, P, 0, TO>(ctx, bs, a, plane, in, f, v, out, w, c);
}

// the marching sweep for 2nd/4th-order operators; returns false if the box
// sizes at this level need another kernel
// whether k_box_march4 runs the sweeps of this level with `extra` bytes of
// LDS per block beyond its two stencil planes (the fused transfers)
template <typename T>
bool march4_applies(const mgx_boxset* bs, int level, size_t extra)
{
    constexpr int V  = march_v<T>();
    constexpr int VM = V > 4 ? V : 4;
    constexpr int ZR = V > 2 ? V : 2;
    const int emax   = bs->maxedge >> level;
    if (emax % VM) return false;
    for (int jb = 0; jb < bs->nbox; jb++)
        if ((bs->edge[jb] >> level) % VM) return false;
    const size_t plane = (size_t)(emax + 2 * ZR) * (emax + 4) * sizeof(T);
    return plane <= 60 * 1024 && (emax * emax) / V <= 1024 && 2 * plane + extra <= 64 * 1024;
}

// LDS bytes of the fused prolongation (three coarse planes) at this level
template <typename T>
size_t march_pro_lds(const mgx_boxset* bs, int level)
{
    const size_t c = (size_t)(bs->maxedge >> (level + 1)) + 1;
    return 3 * c * c * sizeof(T);
}

// rc (JACOBI, output type T, no dots): the stale residual restricted to
// level + 1 inside the sweep, in place of w; pc (the same sweeps): the input
// is in - P(pc), pc the correction at level + 1 (in is not written); false
// if the box sizes need the separate transfer kernels
template <typename T, int TYPE, int MODE, typename TO = T>
bool march(mgx_ctx* ctx, const mgx_boxset* bs, int level, double mjf, const T* in, const T* f, const double* v,
    TO* out, T* w, const double* nrm = nullptr, const MarchDots* dots = nullptr, const double* oinv = nullptr,
    T* rc = nullptr, const T* pc = nullptr, T* rx = nullptr, double rjf = 0.)
{
    constexpr bool mixed = !std::is_same<T, TO>::value;
    static_assert(!mixed || MODE == MARCH_JACOBI, "another output type for the Jacobi sweep only");
    if ((rc || pc) && (MODE != MARCH_JACOBI || mixed || dots || level + 1 >= bs->levels || (rc && pc)))
        return false;
    if (rc && w) return false;
    if (pc && !march4_applies<T>(bs, level, march_pro_lds<T>(bs, level))) return false;
    const int emax = bs->maxedge >> level;
    const int E2   = emax * emax;
    MarchArgs a;
    a.off    = bs->d_off[level];
    a.edge   = bs->d_edge;
    a.origin = bs->d_origin;
    a.level  = level;
    a.g0     = bs->gdim[0] >> level;
    a.g1     = bs->gdim[1] >> level;
    a.g2     = bs->gdim[2] >> level;
    a.lds_w  = emax + 4;
    a.mjf    = mjf;
    a.order  = bs->d_sp_col;
    a.nbox   = bs->nbox;
    if (nrm)
    {
        a.nrm   = nrm;
        a.psi_w = (double*)in;
    }
    Coefs c  = coefs_at(bs, level);
    // rows split into lane groups of march_v points (at least 4)
    constexpr int VM = march_v<T>() > 4 ? march_v<T>() : 4;
    bool by4 = emax % VM == 0;
    for (int jb = 0; by4 && jb < bs->nbox; jb++)
        by4 = ((bs->edge[jb] >> level) % VM) == 0;
    a.oinv = oinv;
    if (mixed && (!by4 || !oinv)) return false;
    if (dots)
    {
        if (!by4) return false;  // the dot products exist in k_box_march4 only
        a.dkr0 = dots->kr0;
        a.da   = dots->a;
        a.dvel = dots->vel;
        a.dout = dots->out;
    }
    if (by4)
    {
        // V points per thread, z ring ZR (see k_box_march4)
        constexpr int V    = march_v<T>();
        constexpr int ZR   = V > 2 ? V : 2;
        a.lds_w            = emax + 2 * ZR;
        const size_t plane = (size_t)a.lds_w * (emax + 4) * sizeof(T);
        const int groups   = E2 / V;
        if (rc)
        {
            if (2 * plane + 2 * (size_t)E2 * sizeof(T) > 64 * 1024 || groups > 1024) return false;
            a.rc   = rc;
            a.offc = bs->d_off[level + 1];
            a.rx   = rx;
            a.rjf  = rjf;
        }
        if (pc)
        {
            const int ecm = (emax >> 1) + 1;
            a.pc          = pc;
            a.cpl         = ecm * ecm;
            a.offc        = bs->d_off[level + 1];
        }
        if (plane <= 60 * 1024 && groups <= 1024)
        {
            if constexpr (TYPE == MGX_LAP_4TH && (MODE == MARCH_JACOBI || MODE == MARCH_HLOC))
            {
                // two lane groups per thread when the autotuner chose so
                constexpr int ti = MODE == MARCH_HLOC ? MGX_TUNE_HLOC_F64
                                                      : (sizeof(T) == 2   ? MGX_TUNE_MARCH_F16
                                                            : sizeof(T) == 4 ? MGX_TUNE_MARCH_F32
                                                                             : MGX_TUNE_MARCH_F64);
                // H psi is tuned for FP64 values only
                const bool two = (MODE != MARCH_HLOC || sizeof(T) == 8) && bs->tune[ti] == 2;
                if (level == 0 && two && groups > 128)
                {
                    if (groups <= 256)
                        march4_launch<T, TYPE, MODE, 128, 2>(ctx, bs, a, plane, in, f, v, out, w, c);
                    else if (groups <= 512)
                        march4_launch<T, TYPE, MODE, 256, 2>(ctx, bs, a, plane, in, f, v, out, w, c);
                    else
                        march4_launch<T, TYPE, MODE, 512, 2>(ctx, bs, a, plane, in, f, v, out, w, c);
                    return true;
                }
            }
            if (groups <= 64)
                march4_launch<T, TYPE, MODE, 64, 1>(ctx, bs, a, plane, in, f, v, out, w, c);
            else if (groups <= 256)
                march4_launch<T, TYPE, MODE, 256, 1>(ctx, bs, a, plane, in, f, v, out, w, c);
            else if (groups <= 512)
                march4_launch<T, TYPE, MODE, 512, 1>(ctx, bs, a, plane, in, f, v, out, w, c);
            else
                march4_launch<T, TYPE, MODE, 1024, 1>(ctx, bs, a, plane, in, f, v, out, w, c);
            return true;
        }
        if (dots) return false;
        a.lds_w = emax + 4;
    }
    if (rc || pc) return false;
    if constexpr (mixed)
        return false;
    else
    {
    if ((size_t)a.lds_w * a.lds_w * sizeof(T) > 60 * 1024) return false;
    if (E2 <= 64)
        march_launch<T, TYPE, MODE, 64, 1>(ctx, bs, a, in, f, v, out, w, c);
    else if (E2 <= 256)
        march_launch<T, TYPE, MODE, 256, 1>(ctx, bs, a, in, f, v, out, w, c);
    else if (E2 <= 512)
        march_launch<T, TYPE, MODE, 256, 2>(ctx, bs, a, in, f, v, out, w, c);
    else if (E2 <= 1024)
        march_launch<T, TYPE, MODE, 256, 4>(ctx, bs, a, in, f, v, out, w, c);
    else if (E2 <= 2048)
        march_launch<T, TYPE, MODE, 256, 8>(ctx, bs, a, in, f, v, out, w, c);
    else
        return false;
    return true;
    }
}

template <typename T>
int box_lap_t(mgx_ctx* ctx, const mgx_boxset* bs, int level, int type, const T* in, T* out)
{
    if (type == MGX_LAP_4TH && march<T, MGX_LAP_4TH, MARCH_LAP>(ctx, bs, level, 0., in, nullptr, nullptr, out, nullptr))
    {
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    if (type == MGX_LAP_2ND && march<T, MGX_LAP_2ND, MARCH_LAP>(ctx, bs, level, 0., in, nullptr, nullptr, out, nullptr))
    {
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    BoxView v = view(bs, level);
    Coefs c   = coefs_at(bs, level);
    dim3 g(tile_grid(v.ntiles)), b(kTile);
    switch (type)
    {
        case MGX_LAP_4TH: hipLaunchKernelGGL((k_box_lap<T, MGX_LAP_4TH>), g, b, 0, ctx->stream, in, out, v, c); break;
        case MGX_LAP_2ND: hipLaunchKernelGGL((k_box_lap<T, MGX_LAP_2ND>), g, b, 0, ctx->stream, in, out, v, c); break;
        case MGX_LAP_MEHRSTELLEN:
            hipLaunchKernelGGL((k_box_lap<T, MGX_LAP_MEHRSTELLEN>), g, b, 0, ctx->stream, in, out, v, c);
            break;
        default: return set_err(ctx, MGX_EINVAL, "LapFactory::createLap() --- option invalid");
    }
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
int box_hloc_t(mgx_ctx* ctx, const mgx_boxset* bs, const T* in, const double* vtot, T* out,
    const double* nrm = nullptr)
{
    if (march<T, MGX_LAP_4TH, MARCH_HLOC>(ctx, bs, 0, 0., in, nullptr, vtot, out, nullptr, nrm))
    {
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    if (nrm) return set_err(ctx, MGX_EINVAL, "fused normalisation needs the marching kernel");
    BoxView v = view(bs, 0);
    Coef4 c4  = coef4(bs->h);
    hipLaunchKernelGGL(k_box_hloc<T>, dim3(tile_grid(v.ntiles)), dim3(kTile), 0, ctx->stream, in, vtot, out, v, c4);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// rc (with rc_done): the stale residual restricted to level + 1 inside the
// sweep when the marching kernel applies (*rc_done = true; w is then not
// written, and rx, if given, receives the coarse level's first sweep
// x = 0 + rjf rc), else w as usual.  pc: the sweep of x - P(pc) (pc at
// level + 1), only where march4_applies says the fused form runs (else an
// error)
template <typename T>
int box_jacobi_t(mgx_ctx* ctx, const mgx_boxset* bs, int level, int type, double jf, const T* x, const T* f, T* xo,
    T* w, const MarchDots* dots = nullptr, bool* dots_done = nullptr, T* rc = nullptr, bool* rc_done = nullptr,
    const T* pc = nullptr, T* rx = nullptr, double rjf = 0.)
{
    const double mjf = -1. * jf;
    if (pc)
    {
        const bool ok = !dots && !w
                        && ((type == MGX_LAP_4TH
                                && march<T, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, level, mjf, x, f, nullptr, xo, nullptr,
                                    nullptr, nullptr, nullptr, nullptr, pc))
                            || (type == MGX_LAP_2ND
                                && march<T, MGX_LAP_2ND, MARCH_JACOBI>(ctx, bs, level, mjf, x, f, nullptr, xo,
                                    nullptr, nullptr, nullptr, nullptr, nullptr, pc)));
        if (!ok) return set_err(ctx, MGX_EINVAL, "fused prolongation needs the marching kernel");
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    if (rc && !dots
        && ((type == MGX_LAP_4TH
                && march<T, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, level, mjf, x, f, nullptr, xo, nullptr, nullptr,
                    nullptr, nullptr, rc, nullptr, rx, rjf))
            || (type == MGX_LAP_2ND
                && march<T, MGX_LAP_2ND, MARCH_JACOBI>(ctx, bs, level, mjf, x, f, nullptr, xo, nullptr, nullptr,
                    nullptr, nullptr, rc, nullptr, rx, rjf))))
    {
        MGX_CHECK_LAUNCH(ctx);
        *rc_done = true;
        return MGX_OK;
    }
    if (dots && type == MGX_LAP_4TH
        && march<T, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, level, mjf, x, f, nullptr, xo, w, nullptr, dots))
    {
        MGX_CHECK_LAUNCH(ctx);
        if (dots_done) *dots_done = true;
        return MGX_OK;
    }
    if (type == MGX_LAP_4TH && march<T, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, level, mjf, x, f, nullptr, xo, w))
    {
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    if (type == MGX_LAP_2ND && march<T, MGX_LAP_2ND, MARCH_JACOBI>(ctx, bs, level, mjf, x, f, nullptr, xo, w))
    {
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    BoxView v = view(bs, level);
    Coefs c   = coefs_at(bs, level);
    dim3 g(tile_grid(v.ntiles)), b(kTile);
    switch (type)
    {
        case MGX_LAP_4TH:
            hipLaunchKernelGGL((k_box_jacobi<T, MGX_LAP_4TH>), g, b, 0, ctx->stream, x, f, xo, w, v, c, mjf);
            break;
        case MGX_LAP_2ND:
            hipLaunchKernelGGL((k_box_jacobi<T, MGX_LAP_2ND>), g, b, 0, ctx->stream, x, f, xo, w, v, c, mjf);
            break;
        case MGX_LAP_MEHRSTELLEN:
            hipLaunchKernelGGL((k_box_jacobi<T, MGX_LAP_MEHRSTELLEN>), g, b, 0, ctx->stream, x, f, xo, w, v, c, mjf);
            break;
        default: return set_err(ctx, MGX_EINVAL, "LapFactory::createLap() --- option invalid");
    }
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
int box_restrict_t(mgx_ctx* ctx, const mgx_boxset* bs, int level, const T* fine, T* coarse)
{
    BoxView vc = view(bs, level + 1);
    // pairs of coarse points per thread when every coarse edge is even
    bool even = ((uintptr_t)fine & 15) == 0;
    for (int e : bs->edge)
        even = even && ((e >> (level + 1)) & 1) == 0;
    if (even)
        hipLaunchKernelGGL(k_box_restrict2p<T>, dim3(tile_grid(vc.ntiles)), dim3(kTile / 2), 0, ctx->stream, fine,
            (const int64_t*)bs->d_off[level], coarse, vc);
    else
        hipLaunchKernelGGL(k_box_restrict2<T>, dim3(tile_grid(vc.ntiles)), dim3(kTile), 0, ctx->stream, fine,
            (const int64_t*)bs->d_off[level], coarse, vc);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T, bool SUB>
int box_prolong_t(mgx_ctx* ctx, const mgx_boxset* bs, int level, const T* coarse, T* fine)
{
    BoxView vc = view(bs, level + 1);
    hipLaunchKernelGGL((k_box_prolong8<T, SUB>), dim3(tile_grid(vc.ntiles)), dim3(kTile), 0, ctx->stream, coarse,
        (const int64_t*)bs->d_off[level], fine, vc);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

template <typename T>
struct BoxMG
{
    std::vector<T*> tmp, w, rc, nv;
};

// the fine-level Jacobi sweep x (T) -> out (FP32), out = x_new * inv[j]
// (the FP16 V-cycle's last sweep); the marching kernel only
template <typename T>
int box_jacobi_out(mgx_ctx* ctx, const mgx_boxset* bs, int type, double jf, const T* x, const T* f, float* out,
    const double* inv, const MarchDots* dots, bool* dots_done)
{
    if constexpr (std::is_same<T, float>::value)
        return set_err(ctx, MGX_EINVAL, "FP32 output from an FP32 V-cycle");
    else
    {
        const double mjf = -1. * jf;
        bool ok          = false;
        if (type == MGX_LAP_4TH)
            ok = march<T, MGX_LAP_4TH, MARCH_JACOBI, float>(ctx, bs, 0, mjf, x, f, nullptr, out, nullptr, nullptr,
                dots, inv);
        else if (type == MGX_LAP_2ND)
            ok = march<T, MGX_LAP_2ND, MARCH_JACOBI, float>(ctx, bs, 0, mjf, x, f, nullptr, out, nullptr, nullptr,
                dots, inv);
        if (!ok) return set_err(ctx, MGX_EINVAL, "FP16 V-cycle needs the marching kernel (edges multiple of 8 per level)");
        MGX_CHECK_LAUNCH(ctx);
        if (dots && dots_done) *dots_done = true;
        return MGX_OK;
    }
}

template <typename T>
int box_mg_setup(mgx_ctx* ctx, const mgx_boxset* bs, int max_levels, BoxMG<T>& B, int slot)
{
    MGX_REQUIRE(ctx, max_levels >= 0 && max_levels < bs->levels, "box set has too few levels");
    size_t total = 0;
    for (int l = 0; l <= max_levels; l++)
    {
        total += 2 * (size_t)bs->off[l][bs->nbox];
        if (l > 0) total += 2 * (size_t)bs->off[l][bs->nbox];
    }
    T* p = (T*)workspace(ctx, slot, total * sizeof(T));
    if (!p) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    for (int l = 0; l <= max_levels; l++)
    {
        const size_t n = bs->off[l][bs->nbox];
        B.tmp.push_back(p);
        p += n;
        B.w.push_back(p);
        p += n;
    }
    for (int l = 1; l <= max_levels; l++)
    {
        const size_t n = bs->off[l][bs->nbox];
        B.rc.push_back(p);
        p += n;
        B.nv.push_back(p);
        p += n;
    }
    return MGX_OK;
}

// Preconditioning<T>::mg, src/Preconditioning.cc:155-216, masks implicit.
// Sweeps ping-pong between v and tmp; every sweep count is even, so the
// result is back in v.
template <typename T>
int box_mg_rec(mgx_ctx* ctx, const mgx_boxset* bs, BoxMG<T>& B, int max_levels, T* v, const T* f, int type,
    int level, bool first_done, const MarchDots* dots = nullptr, bool* dots_done = nullptr, float* last_out = nullptr,
    const double* last_inv = nullptr)
{
    const int ncycl = (level == max_levels) ? 4 : 2;
    double h[3]     = { bs->h[0] * (1 << level), bs->h[1] * (1 << level), bs->h[2] * (1 << level) };
    const double jf = jacobi_factor(type, h);
    // first_done: the zero-start first sweep has been written to tmp; the
    // remaining sweeps ping-pong so that the iterate still ends in v
    T* a = first_done ? B.tmp[level] : v;
    T* b = first_done ? v : B.tmp[level];
    int rc;
    // algorithmic bytes of a fine sweep: x and f read, x written (s each),
    // + w written on the last pre-sweep; the last sweep may write FP32
    // (last_out) and read the previous correction for the Anderson dots
    const double np0 = (double)bs->off[0][bs->nbox];
    const double st  = sizeof(T);
    const double hc[3] = { 2. * h[0], 2. * h[1], 2. * h[2] };
    // the last pre-sweep leaves the stale residual w for the restriction, or
    // restricts it itself (MGX_TUNE_RESTRICT choice 1, where it applies)
    T* rcoarse    = level < max_levels ? B.rc[level] : nullptr;
    bool restricted = false;
    for (int it = first_done ? 1 : 0; it < ncycl; it++)
    {
        T* w          = (it == ncycl - 1 && level < max_levels) ? B.w[level] : nullptr;
        T* rfuse      = (w && bs->tune[MGX_TUNE_RESTRICT] == 1) ? rcoarse : nullptr;
        const int tok = level == 0 ? prof_begin(ctx, MGX_PROF_JACOBI0, (w ? 4. : 3.) * st * np0) : -1;
        // fused, the coarse level's zero-start first sweep is written too
        if ((rc = box_jacobi_t(ctx, bs, level, type, jf, (const T*)a, f, b, w, nullptr, nullptr, rfuse,
                 &restricted, (const T*)nullptr, rfuse ? B.tmp[level + 1] : (T*)nullptr,
                 jacobi_factor(MGX_LAP_2ND, hc))))
            return rc;
        prof_end(ctx, tok);
        // fused: x and f read, x written, the coarse residual and start written
        if (restricted) prof_retag(ctx, tok, MGX_PROF_JACOBI0, (3. + 2. / 8.) * st * np0);
        std::swap(a, b);
    }
    if (level == max_levels) return MGX_OK;
    T* nv   = B.nv[level];
    int tok = -1;
    if (!restricted)
    {
        tok = level == 0 ? prof_begin(ctx, MGX_PROF_TRANSFER, st * np0 * 9. / 8.) : -1;
        if ((rc = box_restrict_t(ctx, bs, level, (const T*)B.w[level], rcoarse))) return rc;
        prof_end(ctx, tok);
    }
    // the coarse level starts from 0: its first sweep is pointwise (written
    // by the fused restriction, else here)
    if (!restricted)
    {
        const int64_t nc = bs->off[level + 1][bs->nbox];
        hipLaunchKernelGGL(k_first_sweep<T>, dim3(nblocks(nc)), dim3(kBlock), 0, ctx->stream,
            (const double*)nullptr, rcoarse, B.tmp[level + 1], nc, jacobi_factor(MGX_LAP_2ND, hc));
        MGX_CHECK_LAUNCH(ctx);
    }
    if ((rc = box_mg_rec(ctx, bs, B, max_levels, nv, (const T*)rcoarse, MGX_LAP_2ND, level + 1, true))) return rc;
    // the coarse correction prolonged inside the first post-sweep
    // (MGX_TUNE_PROLONG choice 1, where the marching kernel runs this level),
    // else by its own kernel into v
    const bool pfuse = bs->tune[MGX_TUNE_PROLONG] == 1 && (type == MGX_LAP_4TH || type == MGX_LAP_2ND)
                       && march4_applies<T>(bs, level, march_pro_lds<T>(bs, level));
    if (!pfuse)
    {
        tok = level == 0 ? prof_begin(ctx, MGX_PROF_TRANSFER, st * np0 * 17. / 8.) : -1;
        if ((rc = box_prolong_t<T, true>(ctx, bs, level, (const T*)nv, v))) return rc;
        prof_end(ctx, tok);
    }
    a = v;
    b = B.tmp[level];
    for (int it = 0; it < 2; it++)
    {
        // fused: x, f and the coarse correction read, x written
        const double pw = (it == 0 && pfuse) ? 1. / 8. : 0.;
        const int tk    = level == 0 ? prof_begin(ctx, MGX_PROF_JACOBI0, (3. + pw) * st * np0) : -1;
        // the last fine sweep (it writes v) may also form the Anderson dots;
        // with last_out it writes FP32 there instead, times last_inv per box
        const MarchDots* dd = (level == 0 && it == 1) ? dots : nullptr;
        bool dd_done        = false;
        if (level == 0 && it == 1 && last_out)
            rc = box_jacobi_out(ctx, bs, type, jf, (const T*)a, f, last_out, last_inv, dd, &dd_done);
        else if (it == 0 && pfuse)
            rc = box_jacobi_t(ctx, bs, level, type, jf, (const T*)a, f, b, (T*)nullptr, nullptr, nullptr,
                (T*)nullptr, nullptr, (const T*)nv);
        else
            rc = box_jacobi_t(ctx, bs, level, type, jf, (const T*)a, f, b, (T*)nullptr, dd, &dd_done);
        if (rc) return rc;
        prof_end(ctx, tk);
        if (level == 0 && it == 1)
        {
            // the previous correction (FP32 from the FP16 V-cycle, else T) is read for the dots
            const double so = last_out ? 4. : st;
            prof_retag(ctx, tk, dd_done ? MGX_PROF_JACOBI0D : MGX_PROF_JACOBI0,
                (2. * st + so + (dd_done ? so : 0.)) * np0);
        }
        if (dd_done && dots_done) *dots_done = true;
        std::swap(a, b);
    }
    return MGX_OK;
}

// OrbitalsPreconditioning::precond_mg on boxes; result left in x (type T);
// out (FP64) written if non-null, else psi += alpha*x if psi non-null.
template <typename T>
int box_precond_t(mgx_ctx* ctx, const mgx_boxset* bs, int type, int max_levels, double gamma, const double* res,
    double* out, double* psi, double alpha, const MarchDots* dots = nullptr, bool* dots_done = nullptr,
    T* xdst = nullptr, bool start_given = false)
{
    MGX_REQUIRE(ctx, type >= 0 && type <= 2, "LapFactory::createLap() --- option invalid");
    MGX_REQUIRE(ctx, gamma >= 0., "gamma must be non-negative");
    const int64_t n = bs->off[0][bs->nbox];
    T* xf           = (T*)workspace(ctx, 2, 2 * n * sizeof(T));
    if (!xf) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    T* x = xdst ? xdst : xf;
    T* f = xf + n;
    BoxMG<T> B;
    int rc = box_mg_setup(ctx, bs, max_levels, B, 3);
    if (rc) return rc;
    // from x0 = 0 the first fine sweep is pointwise and fused with the cast
    const bool zero = gamma == 0.;
    // the residual read (FP64), f and x written (no launch with a given start)
    const int tk = start_given ? -1 : prof_begin(ctx, MGX_PROF_FIRST, (8. + 2. * sizeof(T)) * (double)n);
    // start_given: the residual kernel has written f and the first sweep's x
    // already (box_residual_start)
    MGX_REQUIRE(ctx, zero || !start_given, "a given start is the zero start");
    if (zero && !start_given)
        hipLaunchKernelGGL(k_first_sweep<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, res, f, B.tmp[0], n,
            jacobi_factor(type, bs->h));
    else if (!zero)
        hipLaunchKernelGGL(k_prep<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, res, f, x, n, gamma);
    prof_end(ctx, tk);
    MGX_CHECK_LAUNCH(ctx);
    if ((rc = box_mg_rec(ctx, bs, B, max_levels, x, (const T*)f, type, 0, zero, dots, dots_done))) return rc;
    if (out)
        hipLaunchKernelGGL(k_to_f64<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, (const T*)x, out, n);
    else if (psi)
    {
        // psi read and written, the correction read
        const int tu = prof_begin(ctx, MGX_PROF_UPDATE, (16. + sizeof(T)) * (double)n);
        hipLaunchKernelGGL(k_update<T>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, psi, (const T*)x, n, alpha);
        prof_end(ctx, tu);
    }
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_sumsq(mgx_ctx* ctx, const mgx_boxset* bs, const double* r, double* d);

// The V-cycle with FP16 storage (arithmetic in FP64 as in the other types):
// the residual scaled per box (k_res_scale; dres = vel |R_j|^2 per box, or
// null to form it here), the zero-start or gamma first sweep with the
// scale, FP16 sweeps and transfers on every level, and the last fine sweep
// writing FP32 to xdst with the scale undone (and the Anderson dots, as in
// the other types, from that FP32 output)
int box_precond_h(mgx_ctx* ctx, const mgx_boxset* bs, int type, int max_levels, double gamma, const double* res,
    const double* dres, float* xdst, const MarchDots* dots, bool* dots_done)
{
    using T = _Float16;
    MGX_REQUIRE(ctx, type == MGX_LAP_4TH || type == MGX_LAP_2ND, "FP16 V-cycle: 2nd or 4th order operators only");
    MGX_REQUIRE(ctx, gamma >= 0., "gamma must be non-negative");
    for (int l = 0; l <= max_levels; l++)
        for (int e : bs->edge)
            MGX_REQUIRE(ctx, ((e >> l) % 8) == 0, "FP16 V-cycle: box edges must be multiples of 8 on every level");
    const int64_t n = bs->off[0][bs->nbox];
    T* xf           = (T*)workspace(ctx, 2, 2 * n * sizeof(T));
    double* sc      = (double*)workspace(ctx, 8, 2 * (size_t)bs->nbox * sizeof(double));
    if (!xf || !sc) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    double* inv = sc + bs->nbox;
    T* x        = xf;
    T* f        = xf + n;
    int rc;
    if (!dres)
    {
        if ((rc = box_sumsq(ctx, bs, res, inv))) return rc;
        dres = inv;
    }
    const double vel = bs->h[0] * bs->h[1] * bs->h[2];
    hipLaunchKernelGGL(k_res_scale, dim3((bs->nbox + 255) / 256), dim3(256), 0, ctx->stream, dres,
        (const int64_t*)bs->d_off[0], bs->nbox, vel, sc, inv);
    BoxMG<T> B;
    if ((rc = box_mg_setup(ctx, bs, max_levels, B, 3))) return rc;
    const bool zero = gamma == 0.;
    const int tk    = prof_begin(ctx, MGX_PROF_FIRST, (8. + 2. * sizeof(T)) * (double)n);
    hipLaunchKernelGGL(k_first_sweep_box<T>, dim3(bs->nbox), dim3(256), 0, ctx->stream, res, (const double*)sc, f,
        zero ? B.tmp[0] : x, zero ? jacobi_factor(type, bs->h) : gamma, (const int64_t*)bs->d_off[0], bs->nbox);
    prof_end(ctx, tk);
    MGX_CHECK_LAUNCH(ctx);
    return box_mg_rec(ctx, bs, B, max_levels, x, (const T*)f, type, 0, zero, dots, dots_done, xdst,
        (const double*)inv);
}

// the V-cycle of box_precond_t, its correction (type prec; FP32 for
// MGX_F16) written to xdst (nbox points of the boxes); with kr0 non-null the
// last fine sweep also forms the Anderson dot products of box_anderson_dots
// into d (2 nbox) when its kernel can (*dots_done); dres: the residual's
// per-box vel |R_j|^2 when at hand (FP16 scaling), else null
int box_precond_x(mgx_ctx* ctx, const mgx_boxset* bs, int type, int max_levels, double gamma, int prec,
    const double* res, void* xdst, const void* kr0, double a, double* d, bool* dots_done, const double* dres,
    bool start_given)
{
    MarchDots md{ kr0, a, bs->h[0] * bs->h[1] * bs->h[2], d };
    const MarchDots* dots = kr0 ? &md : nullptr;
    if (dots_done) *dots_done = false;
    if (!xdst) return set_err(ctx, MGX_EINVAL, "no destination for the correction");
    // FP16 storage: the correction comes out in FP32
    if (prec == MGX_F16)
    {
        MGX_REQUIRE(ctx, !start_given, "the FP16 V-cycle forms its own (scaled) start");
        return box_precond_h(ctx, bs, type, max_levels, gamma, res, dres, (float*)xdst, dots, dots_done);
    }
    return prec == MGX_F32 ? box_precond_t<float>(ctx, bs, type, max_levels, gamma, res, nullptr, nullptr, 0., dots,
                                 dots_done, (float*)xdst, start_given)
                           : box_precond_t<double>(ctx, bs, type, max_levels, gamma, res, nullptr, nullptr, 0., dots,
                                 dots_done, (double*)xdst, start_given);
}

int box_precond(mgx_ctx* ctx, const mgx_boxset* bs, int type, int max_levels, double gamma, int prec,
    const double* res, double* out, double* psi, double alpha, bool start_given)
{
    return prec == MGX_F32 ? box_precond_t<float>(ctx, bs, type, max_levels, gamma, res, out, psi, alpha, nullptr,
                                 nullptr, nullptr, start_given)
                           : box_precond_t<double>(ctx, bs, type, max_levels, gamma, res, out, psi, alpha, nullptr,
                                 nullptr, nullptr, start_given);
}

int box_residual_om(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi, const double* theta,
    double* res, double* d, int om, void* fo, void* xo, double jf);

// the buffers box_precond_t keeps its zero start in: f, and the first
// sweep's x (the V-cycle's first temporary of the fine level)
template <typename T>
int precond_start_bufs(mgx_ctx* ctx, const mgx_boxset* bs, int max_levels, T** f, T** x)
{
    const int64_t n = bs->off[0][bs->nbox];
    T* xf           = (T*)workspace(ctx, 2, 2 * n * sizeof(T));
    if (!xf) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    BoxMG<T> B;
    int rc = box_mg_setup(ctx, bs, max_levels, B, 3);
    if (rc) return rc;
    *f = xf + n;
    *x = B.tmp[0];
    return MGX_OK;
}

// R = psi theta - hpsi as the V-cycle's zero start (FP32 or FP64, for a
// box_precond_* call with start_given that follows), with the norms in d; R
// itself is not written.  Needs the cell tables.
int box_residual_start(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi,
    const double* theta, double* d, int prec, int type, int max_levels)
{
    MGX_REQUIRE(ctx, bs->use_cells && (prec == MGX_F32 || prec == MGX_F64),
        "the fused V-cycle start needs the cell tables and an FP32 or FP64 V-cycle");
    const double jf = jacobi_factor(type, bs->h);
    int rc;
    if (prec == MGX_F32)
    {
        float *f, *x;
        if ((rc = precond_start_bufs<float>(ctx, bs, max_levels, &f, &x))) return rc;
        return box_residual_om(ctx, bs, psi, hpsi, theta, nullptr, d, 1, f, x, jf);
    }
    double *f, *x;
    if ((rc = precond_start_bufs<double>(ctx, bs, max_levels, &f, &x))) return rc;
    return box_residual_om(ctx, bs, psi, hpsi, theta, nullptr, d, 2, f, x, jf);
}

bool boxset_has_cells(const mgx_boxset* bs) { return bs->use_cells; }

int box_hloc(mgx_ctx* ctx, const mgx_boxset* bs, int dtype, const void* psi, const double* vtot, void* hpsi)
{
    return dtype == MGX_F32 ? box_hloc_t(ctx, bs, (const float*)psi, vtot, (float*)hpsi)
                            : box_hloc_t(ctx, bs, (const double*)psi, vtot, (double*)hpsi);
}

int box_gram_h_cells(mgx_ctx* ctx, const mgx_boxset* bs, int prec, const double* psi, const double* hpsi, double* S,
    double* H, double* Sv, double* Hv)
{
    const int n      = bs->nbox;
    const size_t pb  = 2 * (size_t)bs->nslot * sizeof(double);
    const size_t sb  = 12 * (size_t)std::max<int64_t>(bs->nilist, 1) * sizeof(double);
    double* part     = (double*)workspace(ctx, 4, pb + sb);
    if (!part) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    double* skin     = part + 2 * bs->nslot;
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(k_fill_b<double>, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, S, nn, 0.);
    hipLaunchKernelGGL(k_fill_b<double>, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, H, nn, 0.);
    {
        // paired loads: the tuned number of waves per cell (1 or 2), else four
        const int nw = bs->tune[prec == MGX_PREC_SPLIT ? MGX_TUNE_GRAM_SPLIT : MGX_TUNE_GRAM_F64] == 2 ? 2 : 1;
        auto kern = prec == MGX_PREC_SPLIT
                        ? (bs->pair_z ? (nw == 2 ? k_cell_gram<MGX_PREC_SPLIT, true, 2>
                                                 : k_cell_gram<MGX_PREC_SPLIT, true, 1>)
                                      : k_cell_gram<MGX_PREC_SPLIT, false, 4>)
                        : (bs->pair_z ? (nw == 2 ? k_cell_gram<MGX_PREC_F64, true, 2> : k_cell_gram<MGX_PREC_F64, true, 1>)
                                      : k_cell_gram<MGX_PREC_F64, false, 4>);
        const int nthr = 64 * (bs->pair_z ? nw : 4);
        // psi and H psi read once (FP64)
        const int tk   = prof_begin(ctx, MGX_PROF_CELLGRAM, 16. * (double)bs->off[0][bs->nbox]);
        hipLaunchKernelGGL(kern, dim3(cell_grid(bs->ncell)), dim3(nthr), 0, ctx->stream, psi, hpsi,
            (const int*)bs->d_cell_ptr, (const CellEnt*)bs->d_cell_ce, (const int64_t*)bs->d_cell_slot, bs->nslot,
            part, bs->ncell);
        prof_end(ctx, tk);
    }
    MGX_CHECK_LAUNCH(ctx);
    if (bs->npairs > 0)
    {
        if (bs->nilist > 0)
        {
            size_t lds        = 2 * (size_t)bs->maxedge * bs->maxedge * sizeof(double);
            const int use_lds = lds <= 64 * 1024;
            if (!use_lds) lds = 0;
            const int tk = prof_begin(ctx, MGX_PROF_SKIN, (double)bs->skin_bytes);
            hipLaunchKernelGGL(k_slab_skin, dim3((6 * bs->nbox + 7) / 8 * 8), dim3(256), lds, ctx->stream, psi,
                (const int64_t*)bs->d_off[0], (const int*)bs->d_edge, (const int*)bs->d_ilist_ptr,
                (const int*)bs->d_ilist, coef4(bs->h), use_lds, skin, (const int*)bs->d_sp_col, bs->nbox);
            prof_end(ctx, tk);
            MGX_CHECK_LAUNCH(ctx);
        }
        const double vel = bs->h[0] * bs->h[1] * bs->h[2];
        hipLaunchKernelGGL(k_pair_reduce, dim3(nblocks(8 * bs->npairs)), dim3(kBlock), 0, ctx->stream,
            (const double*)part, bs->nslot, (const int*)bs->d_pslot_ptr, (const int64_t*)bs->d_pslot,
            (const double*)skin, (const int*)bs->d_pair_ient, (const int*)bs->d_pairs, bs->npairs, n, vel, S, H, Sv,
            Hv);
        MGX_CHECK_LAUNCH(ctx);
    }
    return MGX_OK;
}

// S, H dense (n x n) and, when Sv is not null, their values per pair (Sv, Hv:
// npairs each, the layout the column-tile products read)
int box_gram_h(mgx_ctx* ctx, const mgx_boxset* bs, int prec, const double* psi, const double* hpsi, double* S,
    double* H, double* Sv, double* Hv)
{
    if (bs->use_cells) return box_gram_h_cells(ctx, bs, prec, psi, hpsi, S, H, Sv, Hv);
    const int n = bs->nbox;
    double* part = (double*)workspace(ctx, 4, 3 * std::max<int64_t>(bs->npairs, 1) * sizeof(double));
    if (!part) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(k_fill_b<double>, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, S, nn, 0.);
    hipLaunchKernelGGL(k_fill_b<double>, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, H, nn, 0.);
    if (bs->npairs > 0)
    {
        if (prec == MGX_PREC_SPLIT)
            hipLaunchKernelGGL(k_pair_gram<MGX_PREC_SPLIT>, dim3((unsigned)bs->npairs), dim3(256), 0, ctx->stream,
                psi, hpsi, (const int64_t*)bs->d_off[0], (const int*)bs->d_edge, (const int*)bs->d_pairs,
                bs->npairs, coef4(bs->h), part);
        else
            hipLaunchKernelGGL(k_pair_gram<MGX_PREC_F64>, dim3((unsigned)bs->npairs), dim3(256), 0, ctx->stream,
                psi, hpsi, (const int64_t*)bs->d_off[0], (const int*)bs->d_edge, (const int*)bs->d_pairs,
                bs->npairs, coef4(bs->h), part);
        MGX_CHECK_LAUNCH(ctx);
        const double vel = bs->h[0] * bs->h[1] * bs->h[2];
        hipLaunchKernelGGL(k_pair_scatter, dim3(nblocks(bs->npairs)), dim3(kBlock), 0, ctx->stream,
            (const double*)part, (const int*)bs->d_pairs, bs->npairs, n, vel, S, H, Sv, Hv);
    }
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// Mehrstellen on boxes: H psi (box-truncated), B psi, the Galerkin S, H and
// B-matrix with the operators' one-point reach, and the residual
// (B Psi) theta - H Psi (B psi_a with its reach into the other boxes)
int box_hloc_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* vtot, double* hpsi)
{
    BoxView v = view(bs, 0);
    hipLaunchKernelGGL(k_box_hlocM, dim3(tile_grid(v.ntiles)), dim3(kTile), 0, ctx->stream, psi, vtot, hpsi, v,
        coefm(bs->h));
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_bpsi_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, double* bpsi)
{
    BoxView v = view(bs, 0);
    hipLaunchKernelGGL(k_box_bpsiM, dim3(tile_grid(v.ntiles)), dim3(kTile), 0, ctx->stream, psi, bpsi, v);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_gram_hb_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi, const double* bpsi,
    const double* vtot, double* S, double* H, double* B)
{
    const int n      = bs->nbox;
    const int64_t nn = (int64_t)n * n;
    double* part     = (double*)workspace(ctx, 4, 5 * std::max<int64_t>(bs->npairs, 1) * sizeof(double));
    if (!part) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    for (double* m : { S, H, B })
        hipLaunchKernelGGL(k_fill_b<double>, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, m, nn, 0.);
    if (bs->npairs > 0)
    {
        hipLaunchKernelGGL(k_pair_gram_mehr, dim3((unsigned)bs->npairs), dim3(256), 0, ctx->stream, psi, hpsi, bpsi,
            vtot, (const int64_t*)bs->d_off[0], (const int*)bs->d_edge, (const int*)bs->d_origin,
            (const int*)bs->d_pairs, bs->npairs, bs->gdim[0], bs->gdim[1], bs->gdim[2], coefm(bs->h), part);
        MGX_CHECK_LAUNCH(ctx);
        const double vel = bs->h[0] * bs->h[1] * bs->h[2];
        hipLaunchKernelGGL(k_pair_scatter_mehr, dim3(nblocks(bs->npairs)), dim3(kBlock), 0, ctx->stream,
            (const double*)part, (const int*)bs->d_pairs, bs->npairs, n, vel, S, H, B);
    }
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_residual(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi, const double* theta,
    double* res, double* d);

int box_residual_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* bpsi, const double* hpsi,
    const double* theta, double* res)
{
    int rc = box_residual(ctx, bs, bpsi, hpsi, theta, res, nullptr);
    if (rc) return rc;
    if (bs->nilist > 0)
    {
        hipLaunchKernelGGL(k_resid_skin_mehr, dim3(bs->nbox), dim3(256), 0, ctx->stream, psi, theta, bs->nbox,
            (const int64_t*)bs->d_off[0], (const int*)bs->d_edge, (const int*)bs->d_ilist_ptr,
            (const int*)bs->d_ilist, res);
        MGX_CHECK_LAUNCH(ctx);
    }
    return MGX_OK;
}

// the autotuner (mgx_boxset_autotune): each candidate of each tunable timed
// on scratch buffers of the box set's size; the fastest is kept
int box_autotune(mgx_ctx* ctx, mgx_boxset* bs, double* ms)
{
    const int64_t np = bs->off[0][bs->nbox];
    const int64_t ng = (int64_t)bs->gdim[0] * bs->gdim[1] * bs->gdim[2];
    const int n      = bs->nbox;
    const int64_t nn = (int64_t)n * n;
    const int64_t npr = std::max<int64_t>(bs->npairs, 1);
    double* buf      = nullptr;
    const int64_t total = 3 * np + ng + 2 * nn + 2 * npr;
    if (hipMalloc(&buf, total * sizeof(double)) != hipSuccess)
        return set_err(ctx, MGX_ENOMEM, "autotuner scratch allocation failed");
    hipLaunchKernelGGL(k_fill_b<double>, dim3(nblocks(total)), dim3(kBlock), 0, ctx->stream, buf, total, 0.);
    double *a = buf, *b = a + np, *c = b + np, *v = c + np, *S = v + ng, *H = S + nn, *Sv = H + nn, *Hv = Sv + npr;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const double mjf = -1. * jacobi_factor(MGX_LAP_4TH, bs->h);
    // one tunable: run(cand) launches the kernel with the candidate set;
    // false when the kernel does not apply to this box set
    auto tune1 = [&](int ti, int ncand, auto&& run) {
        double best = 1e300;
        int bestc    = 1;
        for (int cand = 1; cand <= ncand; cand++)
        {
            bs->tune[ti] = cand;
            double t     = -1.;
            if (run())
            {
                float acc[3];
                for (int r = 0; r < 3; r++)
                {
                    (void)hipEventRecord(e0, ctx->stream);
                    run();
                    (void)hipEventRecord(e1, ctx->stream);
                    (void)hipEventSynchronize(e1);
                    (void)hipEventElapsedTime(&acc[r], e0, e1);
                }
                std::sort(acc, acc + 3);
                t = acc[1];
                if (t < best)
                {
                    best  = t;
                    bestc = cand;
                }
            }
            if (ms) ms[2 * ti + cand - 1] = t;
        }
        bs->tune[ti] = bestc;
    };
    tune1(MGX_TUNE_MARCH_F16, 2, [&]() {
        return march<_Float16, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, 0, mjf, (const _Float16*)a, (const _Float16*)b,
            nullptr, (_Float16*)c, nullptr);
    });
    tune1(MGX_TUNE_MARCH_F32, 2, [&]() {
        return march<float, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, 0, mjf, (const float*)a, (const float*)b, nullptr,
            (float*)c, nullptr);
    });
    tune1(MGX_TUNE_MARCH_F64, 2, [&]() {
        return march<double, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, 0, mjf, a, b, nullptr, c, nullptr);
    });
    tune1(MGX_TUNE_HLOC_F64, 2, [&]() {
        return march<double, MGX_LAP_4TH, MARCH_HLOC>(ctx, bs, 0, 0., a, nullptr, v, c, nullptr);
    });
    // the FP32 pre-sweep with the restriction fused, or followed by it
    tune1(MGX_TUNE_RESTRICT, 2, [&]() {
        if (bs->levels < 2) return false;
        float *x = (float*)a, *w = x + np, *fv = (float*)b, *rcv = fv + np, *xo = (float*)c;
        if (bs->tune[MGX_TUNE_RESTRICT] == 1)
            return march<float, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, 0, mjf, x, fv, nullptr, xo, nullptr, nullptr,
                nullptr, nullptr, rcv);
        return march<float, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, 0, mjf, x, fv, nullptr, xo, w)
               && box_restrict_t<float>(ctx, bs, 0, w, rcv) == MGX_OK;
    });
    // the FP32 post-sweep with the prolongation fused, or after it
    tune1(MGX_TUNE_PROLONG, 2, [&]() {
        if (bs->levels < 2) return false;
        float *x = (float*)a, *fv = (float*)b, *pcv = fv + np, *xo = (float*)c;
        if (bs->tune[MGX_TUNE_PROLONG] == 1)
            return march<float, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, 0, mjf, x, fv, nullptr, xo, nullptr, nullptr,
                nullptr, nullptr, nullptr, pcv);
        return box_prolong_t<float, true>(ctx, bs, 0, pcv, x) == MGX_OK
               && march<float, MGX_LAP_4TH, MARCH_JACOBI>(ctx, bs, 0, mjf, x, fv, nullptr, xo, nullptr);
    });
    for (int prec : { MGX_PREC_F64, MGX_PREC_SPLIT })
        tune1(prec == MGX_PREC_F64 ? MGX_TUNE_GRAM_F64 : MGX_TUNE_GRAM_SPLIT, 2, [&]() {
            return bs->use_cells && bs->pair_z && box_gram_h_cells(ctx, bs, prec, a, b, S, H, Sv, Hv) == MGX_OK;
        });
    const hipError_t err = hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    if (err != hipSuccess) return set_err(ctx, MGX_EHIP, std::string("autotuner: ") + hipGetErrorString(err));
    bs->tuned = true;
    return MGX_OK;
}

int box_sumsq(mgx_ctx* ctx, const mgx_boxset* bs, const double* r, double* d);

// the cell residual kernel for output mode OM (R, or the V-cycle's zero start)
template <int OM>
void cell_residual_launch(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi,
    const double* theta, double* res, double* esq, void* fo, void* xo, double jf)
{
    auto kern = bs->pair_z ? k_cell_residual<true, OM> : k_cell_residual<false, OM>;
    hipLaunchKernelGGL(kern, dim3(cell_grid(bs->ncell)), dim3(256), 0, ctx->stream, psi, hpsi, theta, bs->nbox,
        (const int*)bs->d_cell_ptr, (const CellEnt*)bs->d_cell_ce, res, esq, bs->ncell, fo, xo, jf);
}

// R = psi theta - hpsi per box; with d: also d_j = vel |R_j|^2 (computed in
// the cell kernel when the cell tables exist, else by k_box_sumsq).  om 1 / 2
// (cell tables only): instead of R the kernel writes the FP32 / FP64 zero
// start of the V-cycle, f = (T)R to fo and x = 0 + (T)(jf f) to xo
int box_residual_om(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi, const double* theta,
    double* res, double* d, int om, void* fo, void* xo, double jf)
{
    if (bs->use_cells)
    {
        double* esq = nullptr;
        if (d)
        {
            // its own slot: the V-cycle's buffers (slot 3) may hold the start
            esq = (double*)workspace(ctx, 9, (size_t)std::max(bs->nent, 1) * sizeof(double));
            if (!esq) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
        }
        // psi, H psi read; R written (FP64), or f and x (2 x the V-cycle's type)
        const double wb = om == 0 ? 8. : (om == 1 ? 8. : 16.);
        const int tk    = prof_begin(ctx, MGX_PROF_RESIDK, (16. + wb) * (double)bs->off[0][bs->nbox]);
        if (om == 1)
            cell_residual_launch<1>(ctx, bs, psi, hpsi, theta, res, esq, fo, xo, jf);
        else if (om == 2)
            cell_residual_launch<2>(ctx, bs, psi, hpsi, theta, res, esq, fo, xo, jf);
        else
            cell_residual_launch<0>(ctx, bs, psi, hpsi, theta, res, esq, nullptr, nullptr, 0.);
        prof_end(ctx, tk);
        if (d)
        {
            const double vel = bs->h[0] * bs->h[1] * bs->h[2];
            hipLaunchKernelGGL(k_box_from_entries, dim3(bs->nbox), dim3(64), 0, ctx->stream, (const double*)esq,
                (const int*)bs->d_bent_ptr, (const int*)bs->d_bent, bs->nbox, vel, d);
        }
        MGX_CHECK_LAUNCH(ctx);
        return MGX_OK;
    }
    MGX_REQUIRE(ctx, om == 0, "the fused V-cycle start needs the cell tables");
    BoxView v = view(bs, 0);
    hipLaunchKernelGGL(k_box_residual, dim3(tile_grid(v.ntiles)), dim3(kTile), 0, ctx->stream, psi, hpsi, theta,
        bs->nbox, (const int*)bs->d_nbr_ptr, (const int*)bs->d_nbr, v, res);
    MGX_CHECK_LAUNCH(ctx);
    return d ? box_sumsq(ctx, bs, res, d) : MGX_OK;
}

int box_residual(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi, const double* theta,
    double* res, double* d)
{
    return box_residual_om(ctx, bs, psi, hpsi, theta, res, d, 0, nullptr, nullptr, 0.);
}

// rho on the level-0 grid (x slowest, z fastest) from the boxes' psi and
// the density matrix X (n x n, column-major): zeroed, then written cell by
// cell; *charge = vel * sum rho (Rho::computeTotalCharge,
// src/Rho.cc:122-153).  With nel > 0, rho is rescaled to nel electrons
// (Rho::rescaleTotalCharge, src/Rho.cc:155-208), which fails with
// MGX_ENUMERIC, rho left unscaled, when the factor is off 1 by more than 2%.
int box_density(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* X, int n, double nel,
    double* rho, double* charge)
{
    if (!bs->use_cells)
        return set_err(ctx, MGX_EINVAL,
            "the density needs the cell tables (grid edges multiples of 8, at most 32 boxes per cell)");
    if (n != bs->nbox) return set_err(ctx, MGX_EINVAL, "density matrix size differs from the number of boxes");
    const int64_t ng = (int64_t)bs->gdim[0] * bs->gdim[1] * bs->gdim[2];
    double* cs       = (double*)workspace(ctx, 6, (bs->ncell + 8) * sizeof(double));
    if (!cs) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    hipLaunchKernelGGL(k_fill_b<double>, dim3(nblocks(ng)), dim3(kBlock), 0, ctx->stream, rho, ng, 0.);
    hipLaunchKernelGGL(bs->pair_z ? k_cell_density<true> : k_cell_density<false>, dim3(cell_grid(bs->ncell)),
        dim3(256), 0, ctx->stream, psi, X, n, (const int*)bs->d_origin, bs->gdim[0], bs->gdim[1], bs->gdim[2],
        (const int*)bs->d_cell_ptr, (const CellEnt*)bs->d_cell_ce, rho, cs, bs->ncell);
    const double vel = bs->h[0] * bs->h[1] * bs->h[2];
    hipLaunchKernelGGL(k_sum_scaled, dim3(1), dim3(256), 0, ctx->stream, (const double*)cs, bs->ncell, vel,
        cs + bs->ncell);
    MGX_CHECK_LAUNCH(ctx);
    double tc = 0.;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(&tc, cs + bs->ncell, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (charge) *charge = tc;
    if (nel > 0. && tc > 0.)
    {
        const double t1 = nel / tc;
        if (std::fabs(t1 - 1.) > 0.02)
            return set_err(ctx, MGX_ENUMERIC, "density charge differs from the electron count by more than 2%");
        hipLaunchKernelGGL(k_scale_d, dim3(nblocks(ng)), dim3(kBlock), 0, ctx->stream, rho, ng, t1);
        MGX_CHECK_LAUNCH(ctx);
    }
    return MGX_OK;
}

// Anderson m = 1 inner products, weighted by diag(S^-1): out[0] =
// <f - f0, f - f0>, out[1] = <f - f0, f> (device), f = a*kr, f0 = a*kr0;
// kr, kr0 of type prec (MGX_F32 / MGX_F64).  d: 2 n scratch.
// the S^-1-weighted totals of per-box dots already in d (see box_precond_x)
int box_anderson_dots_total(mgx_ctx* ctx, const mgx_boxset* bs, const double* Sinv, const double* d, double* out)
{
    hipLaunchKernelGGL(k_diag_wsum2, dim3(1), dim3(256), 0, ctx->stream, d, Sinv, bs->nbox, out);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_anderson_dots(mgx_ctx* ctx, const mgx_boxset* bs, int prec, const void* kr, double a, const void* kr0,
    const double* Sinv, double* d, double* out)
{
    const double vel = bs->h[0] * bs->h[1] * bs->h[2];
    if (prec == MGX_F32)
        hipLaunchKernelGGL(k_box_and_dots<float>, dim3(bs->nbox), dim3(256), 0, ctx->stream, (const float*)kr, a,
            (const float*)kr0, (const int64_t*)bs->d_off[0], bs->nbox, vel, d);
    else
        hipLaunchKernelGGL(k_box_and_dots<double>, dim3(bs->nbox), dim3(256), 0, ctx->stream, (const double*)kr, a,
            (const double*)kr0, (const int64_t*)bs->d_off[0], bs->nbox, vel, d);
    hipLaunchKernelGGL(k_diag_wsum2, dim3(1), dim3(256), 0, ctx->stream, (const double*)d, Sinv, bs->nbox, out);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// Anderson update into xo (see k_box_and_update_sq) + the norms d of the new
// x; theta read from device memory (null: 0); kr0_out, if non-null, receives
// a copy of kr (the next step's history when kr's buffer is not kept)
int box_anderson_update_sq(mgx_ctx* ctx, const mgx_boxset* bs, int prec, const double* x, double* xo, const void* kr,
    double a, const void* kr0, void* kr0_out, const double* th, double beta, double* d)
{
    const double vel = bs->h[0] * bs->h[1] * bs->h[2];
    const int64_t* off = (const int64_t*)bs->d_off[0];
    if (prec == MGX_F32)
        hipLaunchKernelGGL(k_box_and_update_sq<float>, dim3(bs->nbox), dim3(256), 0, ctx->stream, x, xo,
            (const float*)kr, a, (const float*)kr0, (float*)kr0_out, off, bs->nbox, th, beta, vel, d);
    else
        hipLaunchKernelGGL(k_box_and_update_sq<double>, dim3(bs->nbox), dim3(256), 0, ctx->stream, x, xo,
            (const double*)kr, a, (const double*)kr0, (double*)kr0_out, off, bs->nbox, th, beta, vel, d);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// box_anderson_update_sq with the normalisation applied in the same kernel
// (boxes of at most 32^3 points); returns MGX_EINVAL for larger boxes
int box_anderson_update_nrm(mgx_ctx* ctx, const mgx_boxset* bs, int prec, const double* x, double* xo,
    const void* kr, double a, const void* kr0, void* kr0_out, const double* th, double beta, double* d)
{
    if (bs->maxedge > 32) return set_err(ctx, MGX_EINVAL, "in-register normalisation takes boxes up to 32^3");
    const double vel = bs->h[0] * bs->h[1] * bs->h[2];
    const int64_t* off = (const int64_t*)bs->d_off[0];
    if (prec == MGX_F32)
        hipLaunchKernelGGL((k_box_and_update_nrm<float, 32>), dim3(bs->nbox), dim3(1024), 0, ctx->stream, x, xo,
            (const float*)kr, a, (const float*)kr0, (float*)kr0_out, off, bs->nbox, th, beta, vel, d);
    else
        hipLaunchKernelGGL((k_box_and_update_nrm<double, 32>), dim3(bs->nbox), dim3(1024), 0, ctx->stream, x, xo,
            (const double*)kr, a, (const double*)kr0, (double*)kr0_out, off, bs->nbox, th, beta, vel, d);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_psi_step_sq(mgx_ctx* ctx, const mgx_boxset* bs, int prec, double* psi, const void* kr, double a, double* d)
{
    const double vel = bs->h[0] * bs->h[1] * bs->h[2];
    const int64_t* off = (const int64_t*)bs->d_off[0];
    if (prec == MGX_F32)
        hipLaunchKernelGGL(k_box_psi_step_sq<float>, dim3(bs->nbox), dim3(256), 0, ctx->stream, psi, (const float*)kr,
            a, off, bs->nbox, vel, d);
    else
        hipLaunchKernelGGL(k_box_psi_step_sq<double>, dim3(bs->nbox), dim3(256), 0, ctx->stream, psi,
            (const double*)kr, a, off, bs->nbox, vel, d);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_sumsq(mgx_ctx* ctx, const mgx_boxset* bs, const double* r, double* d)
{
    const double vel = bs->h[0] * bs->h[1] * bs->h[2];
    hipLaunchKernelGGL(k_box_sumsq, dim3(bs->nbox), dim3(256), 0, ctx->stream, r, (const int64_t*)bs->d_off[0],
        bs->nbox, vel, d);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_norm(mgx_ctx* ctx, const mgx_boxset* bs, const double* d, const double* Sinv, double* dev_out)
{
    hipLaunchKernelGGL(k_norm, dim3(1), dim3(256), 0, ctx->stream, d, Sinv, bs->nbox, dev_out);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// Hpsi of the FP64 orbitals with the pending normalisation d (vel*|psi_j|^2
// of the unnormalised psi) applied on the way, psi rewritten normalised
int box_hloc_normalize(mgx_ctx* ctx, const mgx_boxset* bs, double* psi, const double* vtot, double* hpsi,
    const double* d)
{
    return box_hloc_t(ctx, bs, (const double*)psi, vtot, hpsi, d);
}

int box_scale(mgx_ctx* ctx, const mgx_boxset* bs, double* psi, const double* d)
{
    BoxView v = view(bs, 0);
    hipLaunchKernelGGL(k_box_scale, dim3(tile_grid(v.ntiles)), dim3(kTile), 0, ctx->stream, psi,
        (const int64_t*)bs->d_off[0], (const int*)bs->d_tile[0], bs->ntiles[0], (const int*)bs->d_edge, d);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int box_normalize(mgx_ctx* ctx, const mgx_boxset* bs, double* psi, double* d)
{
    int rc = box_sumsq(ctx, bs, psi, d);
    if (rc) return rc;
    BoxView v = view(bs, 0);
    hipLaunchKernelGGL(k_box_scale, dim3(tile_grid(v.ntiles)), dim3(kTile), 0, ctx->stream, psi,
        (const int64_t*)bs->d_off[0], (const int*)bs->d_tile[0], bs->ntiles[0], (const int*)bs->d_edge,
        (const double*)d);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int64_t box_size(const mgx_boxset* bs, int level) { return bs->off[level][bs->nbox]; }

PairTiles boxset_tiles(const mgx_boxset* bs)
{
    PairTiles t;
    t.ntile = bs->ntile;
    t.ptr   = bs->d_sp_ptr;
    t.col   = bs->d_sp_col;
    t.k     = bs->d_sp_k;
    t.e     = bs->d_sp_e;
    return t;
}

int64_t boxset_npairs(const mgx_boxset* bs) { return bs->npairs; }

int boxset_maxedge(const mgx_boxset* bs) { return bs->maxedge; }

bool boxset_uniform_edge(const mgx_boxset* bs, int* edge)
{
    *edge = bs->edge.empty() ? 0 : bs->edge[0];
    for (int e : bs->edge)
        if (e != *edge) return false;
    return true;
}

void boxset_geom(const mgx_boxset* bs, int gdim[3], double h[3], int* nbox)
{
    for (int d = 0; d < 3; d++)
    {
        gdim[d] = bs->gdim[d];
        h[d]    = bs->h[d];
    }
    *nbox = bs->nbox;
}

}  // namespace mgx

using namespace mgx;

extern "C" {

int mgx_boxset_create(mgx_ctx* ctx, const mgx_boxes_t* b, int levels, mgx_boxset** out)
{
    MGX_REQUIRE(ctx, b && out && b->origin && b->edge && b->nbox > 0, "bad box table");
    MGX_REQUIRE(ctx, levels >= 1 && levels <= 8, "levels must be in [1,8]");
    const int m = 1 << (levels - 1);
    for (int d = 0; d < 3; d++)
        MGX_REQUIRE(ctx, b->gdim[d] > 0 && b->gdim[d] % m == 0 && b->h[d] > 0., "bad global grid");
    for (int j = 0; j < b->nbox; j++)
    {
        MGX_REQUIRE(ctx, b->edge[j] >= m && b->edge[j] % m == 0 && b->edge[j] <= 1024, "box edge not aligned");
        for (int d = 0; d < 3; d++)
        {
            MGX_REQUIRE(ctx, b->origin[3 * j + d] % m == 0, "box origin not aligned");
            MGX_REQUIRE(ctx, b->origin[3 * j + d] >= 0 && b->origin[3 * j + d] < b->gdim[d], "box origin outside grid");
            MGX_REQUIRE(ctx, 2 * b->edge[j] + 4 <= b->gdim[d], "box edge larger than half the grid minus 2");
        }
    }
    mgx_boxset* s = new mgx_boxset();
    s->ctx        = ctx;
    s->nbox       = b->nbox;
    s->levels     = levels;
    for (int d = 0; d < 3; d++)
    {
        s->gdim[d] = b->gdim[d];
        s->h[d]    = b->h[d];
    }
    s->origin.assign(b->origin, b->origin + 3 * b->nbox);
    s->edge.assign(b->edge, b->edge + b->nbox);
    s->maxedge = *std::max_element(s->edge.begin(), s->edge.end());
    auto fail  = [&](const char* what) {
        mgx_boxset_destroy(s);
        return set_err(ctx, MGX_EHIP, what);
    };
    s->off.resize(levels);
    s->d_off.assign(levels, nullptr);
    s->ntiles.assign(levels, 0);
    s->d_tile.assign(levels, nullptr);
    for (int l = 0; l < levels; l++)
    {
        auto& o = s->off[l];
        o.resize(s->nbox + 1);
        int64_t acc = 0;
        std::vector<int> tiles;
        for (int j = 0; j < s->nbox; j++)
        {
            o[j]        = acc;
            const int e = s->edge[j] >> l;
            const int n = e * e * e;
            acc += n;
            for (int st = 0; st < n; st += kTile)
            {
                tiles.push_back(j);
                tiles.push_back(st);
            }
        }
        o[s->nbox]   = acc;
        s->ntiles[l] = (int)(tiles.size() / 2);
        if (hipMalloc(&s->d_off[l], sizeof(int64_t) * (s->nbox + 1)) != hipSuccess) return fail("hipMalloc offsets");
        if (hipMalloc(&s->d_tile[l], sizeof(int) * tiles.size()) != hipSuccess) return fail("hipMalloc tiles");
        hipMemcpy(s->d_off[l], o.data(), sizeof(int64_t) * (s->nbox + 1), hipMemcpyHostToDevice);
        hipMemcpy(s->d_tile[l], tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice);
    }
    if (hipMalloc(&s->d_origin, sizeof(int) * 3 * s->nbox) != hipSuccess) return fail("hipMalloc origin");
    if (hipMalloc(&s->d_edge, sizeof(int) * s->nbox) != hipSuccess) return fail("hipMalloc edge");
    hipMemcpy(s->d_origin, s->origin.data(), sizeof(int) * 3 * s->nbox, hipMemcpyHostToDevice);
    hipMemcpy(s->d_edge, s->edge.data(), sizeof(int) * s->nbox, hipMemcpyHostToDevice);

    // overlap tables (level 0), sorted by (i, j)
    std::vector<int> pairs;
    std::vector<std::vector<int>> nb(s->nbox);
    for (int i = 0; i < s->nbox; i++)
        for (int j = i; j < s->nbox; j++)
        {
            int a[3], bb[3], sh[3];
            bool inter = true;
            for (int d = 0; d < 3 && inter; d++)
                inter = ext_shift1d(s->origin[3 * i + d], s->edge[i], s->origin[3 * j + d], s->edge[j], s->gdim[d],
                    &sh[d]);
            if (!inter) continue;
            bool ok = true;
            for (int d = 0; d < 3; d++)
            {
                // overlap in i-local coordinates: x in [0,ei) with x+sh in [0,ej)
                a[d]  = std::max(0, -sh[d]);
                bb[d] = std::min(s->edge[i], s->edge[j] - sh[d]);
                if (a[d] >= bb[d]) ok = false;
            }
            pairs.insert(pairs.end(), { i, j, a[0], a[1], a[2], bb[0], bb[1], bb[2], sh[0], sh[1], sh[2] });
            if (!ok) continue;
            // neighbour records in j-local coordinates for box j (k = i) and
            // in i-local coordinates for box i (k = j)
            nb[i].insert(nb[i].end(), { j, a[0], a[1], a[2], bb[0], bb[1], bb[2], sh[0], sh[1], sh[2] });
            if (i != j)
                nb[j].insert(nb[j].end(), { i, a[0] + sh[0], a[1] + sh[1], a[2] + sh[2], bb[0] + sh[0],
                                              bb[1] + sh[1], bb[2] + sh[2], -sh[0], -sh[1], -sh[2] });
        }
    s->npairs = (int64_t)(pairs.size() / 11);
    {
        // interaction lists for the skin terms (other boxes only)
        std::vector<std::vector<int>> il(s->nbox);
        std::vector<int> pient(2 * std::max<int64_t>(s->npairs, 1), -1);
        for (int64_t q = 0; q < s->npairs; q++)
        {
            const int* P = &pairs[11 * q];
            const int i = P[0], j = P[1];
            if (i == j) continue;
            // entry of i in j's list: sh maps j-local -> i-local = -sh_ij
            pient[2 * q] = (int)(il[j].size() / 4);
            il[j].insert(il[j].end(), { i, -P[8], -P[9], -P[10] });
            pient[2 * q + 1] = (int)(il[i].size() / 4);
            il[i].insert(il[i].end(), { j, P[8], P[9], P[10] });
        }
        std::vector<int> iptr(s->nbox + 1, 0), iflat;
        for (int j = 0; j < s->nbox; j++)
        {
            iptr[j] = (int)(iflat.size() / 4);
            iflat.insert(iflat.end(), il[j].begin(), il[j].end());
        }
        iptr[s->nbox] = (int)(iflat.size() / 4);
        // the skin's algorithmic bytes, by k_slab_skin's own tests
        for (int j = 0; j < s->nbox; j++)
        {
            const int ej = s->edge[j];
            for (int f = 0; f < 6; f++)
            {
                const int d = f >> 1, side = f & 1, u = (d + 1) % 3, w = (d + 2) % 3;
                const int Xd1 = side == 0 ? -1 : ej, Xd2 = side == 0 ? -2 : ej + 1;
                bool any = false;
                for (size_t q = 0; q < il[j].size(); q += 4)
                {
                    const int i = il[j][q];
                    if (i >= j) continue;
                    const int ei = s->edge[i];
                    const int sh[3] = { il[j][q + 1], il[j][q + 2], il[j][q + 3] };
                    const int nu  = std::min(ej, ei - sh[u]) - std::max(0, -sh[u]);
                    const int nww = std::min(ej, ei - sh[w]) - std::max(0, -sh[w]);
                    if (nu <= 0 || nww <= 0) continue;
                    const int xd1 = Xd1 + sh[d], xd2 = Xd2 + sh[d];
                    const int layers = (xd1 >= 0 && xd1 < ei ? 1 : 0) + (xd2 >= 0 && xd2 < ei ? 1 : 0);
                    if (!layers) continue;
                    any = true;
                    s->skin_bytes += 8 * (int64_t)nu * nww * layers;
                }
                if (any) s->skin_bytes += 2 * 8 * (int64_t)ej * ej;
            }
        }
        // pair entries are local to each box's list: make them global
        for (int64_t q = 0; q < s->npairs; q++)
        {
            const int i = pairs[11 * q], j = pairs[11 * q + 1];
            if (i == j) continue;
            pient[2 * q] += iptr[j];
            pient[2 * q + 1] += iptr[i];
        }
        s->nilist = iptr[s->nbox];
        if (hipMalloc(&s->d_ilist_ptr, sizeof(int) * iptr.size()) != hipSuccess
            || hipMalloc(&s->d_ilist, sizeof(int) * std::max<size_t>(iflat.size(), 1)) != hipSuccess
            || hipMalloc(&s->d_pair_ient, sizeof(int) * pient.size()) != hipSuccess)
            return fail("hipMalloc interaction lists");
        hipMemcpy(s->d_ilist_ptr, iptr.data(), sizeof(int) * iptr.size(), hipMemcpyHostToDevice);
        if (!iflat.empty()) hipMemcpy(s->d_ilist, iflat.data(), sizeof(int) * iflat.size(), hipMemcpyHostToDevice);
        hipMemcpy(s->d_pair_ient, pient.data(), sizeof(int) * pient.size(), hipMemcpyHostToDevice);
    }
    {
        // column tiles for the pair-sparse products: boxes sorted along a
        // Morton curve of their centres, kSpTile consecutive ones per tile,
        // so that a tile's columns share most of their rows
        auto spread = [](uint64_t v) {
            uint64_t r = 0;
            for (int b = 0; b < 21; b++)
                r |= ((v >> b) & 1ull) << (3 * b);
            return r;
        };
        std::vector<std::pair<uint64_t, int>> key(s->nbox);
        for (int j = 0; j < s->nbox; j++)
        {
            uint64_t c[3];
            for (int d = 0; d < 3; d++)
                c[d] = (uint64_t)((s->origin[3 * j + d] + s->edge[j] / 2) % s->gdim[d]);
            key[j] = { (spread(c[0]) << 2) | (spread(c[1]) << 1) | spread(c[2]), j };
        }
        std::sort(key.begin(), key.end());
        std::vector<std::vector<std::pair<int, int>>> colk(s->nbox);  // (row k, pair)
        for (int64_t q = 0; q < s->npairs; q++)
        {
            const int i = pairs[11 * q], j = pairs[11 * q + 1];
            colk[j].push_back({ i, (int)q });
            if (i != j) colk[i].push_back({ j, (int)q });
        }
        s->ntile = (s->nbox + kSpTile - 1) / kSpTile;
        std::vector<int> tptr(s->ntile + 1, 0), tcol((size_t)s->ntile * kSpTile, -1), tk, te;
        std::vector<int> pos(s->nbox, -1);
        for (int t = 0; t < s->ntile; t++)
        {
            tptr[t] = (int)tk.size();
            std::vector<int> rows;
            for (int q = 0; q < kSpTile && t * kSpTile + q < s->nbox; q++)
            {
                const int r              = key[t * kSpTile + q].second;
                tcol[t * kSpTile + q]    = r;
                for (auto& kp : colk[r])
                    if (pos[kp.first] < 0)
                    {
                        pos[kp.first] = 0;
                        rows.push_back(kp.first);
                    }
            }
            std::sort(rows.begin(), rows.end());
            const size_t base = tk.size();
            for (size_t u = 0; u < rows.size(); u++)
                pos[rows[u]] = (int)u;
            tk.insert(tk.end(), rows.begin(), rows.end());
            te.resize(te.size() + rows.size() * kSpTile, -1);
            for (int q = 0; q < kSpTile && t * kSpTile + q < s->nbox; q++)
                for (auto& kp : colk[tcol[t * kSpTile + q]])
                    te[(base + pos[kp.first]) * kSpTile + q] = kp.second;
            for (int r : rows)
                pos[r] = -1;
        }
        tptr[s->ntile] = (int)tk.size();
        if (hipMalloc(&s->d_sp_ptr, sizeof(int) * tptr.size()) != hipSuccess
            || hipMalloc(&s->d_sp_col, sizeof(int) * tcol.size()) != hipSuccess
            || hipMalloc(&s->d_sp_k, sizeof(int) * std::max<size_t>(tk.size(), 1)) != hipSuccess
            || hipMalloc(&s->d_sp_e, sizeof(int) * std::max<size_t>(te.size(), 1)) != hipSuccess)
            return fail("hipMalloc column tiles");
        hipMemcpy(s->d_sp_ptr, tptr.data(), sizeof(int) * tptr.size(), hipMemcpyHostToDevice);
        hipMemcpy(s->d_sp_col, tcol.data(), sizeof(int) * tcol.size(), hipMemcpyHostToDevice);
        if (!tk.empty())
        {
            hipMemcpy(s->d_sp_k, tk.data(), sizeof(int) * tk.size(), hipMemcpyHostToDevice);
            hipMemcpy(s->d_sp_e, te.data(), sizeof(int) * te.size(), hipMemcpyHostToDevice);
        }
    }
    std::vector<int> ptr(s->nbox + 1, 0), flat;
    for (int j = 0; j < s->nbox; j++)
    {
        // sort records of box j by k
        const int cnt = (int)(nb[j].size() / 10);
        std::vector<int> idx(cnt);
        for (int q = 0; q < cnt; q++)
            idx[q] = q;
        std::sort(idx.begin(), idx.end(), [&](int u, int w) { return nb[j][10 * u] < nb[j][10 * w]; });
        ptr[j] = (int)(flat.size() / 10);
        for (int q : idx)
            flat.insert(flat.end(), nb[j].begin() + 10 * q, nb[j].begin() + 10 * q + 10);
    }
    ptr[s->nbox] = (int)(flat.size() / 10);
    s->nnbr      = ptr[s->nbox];
    if (hipMalloc(&s->d_pairs, sizeof(int) * std::max<size_t>(pairs.size(), 1)) != hipSuccess)
        return fail("hipMalloc pairs");
    if (hipMalloc(&s->d_nbr_ptr, sizeof(int) * ptr.size()) != hipSuccess) return fail("hipMalloc nbr ptr");
    if (hipMalloc(&s->d_nbr, sizeof(int) * std::max<size_t>(flat.size(), 1)) != hipSuccess) return fail("hipMalloc nbr");
    if (!pairs.empty()) hipMemcpy(s->d_pairs, pairs.data(), sizeof(int) * pairs.size(), hipMemcpyHostToDevice);
    hipMemcpy(s->d_nbr_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice);
    if (!flat.empty()) hipMemcpy(s->d_nbr, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice);
    // spatial cells (level 0)
    bool cells_ok = true;
    for (int d = 0; d < 3; d++)
        cells_ok = cells_ok && (s->gdim[d] % kCell == 0);
    if (cells_ok)
    {
        const int C0 = s->gdim[0] / kCell, C1 = s->gdim[1] / kCell, C2 = s->gdim[2] / kCell;
        std::vector<std::vector<int>> cl((size_t)C0 * C1 * C2);
        for (int j = 0; j < s->nbox; j++)
        {
            const int e = s->edge[j];
            int c_lo[3], c_hi[3];
            for (int d = 0; d < 3; d++)
            {
                c_lo[d] = s->origin[3 * j + d] / kCell;
                c_hi[d] = (s->origin[3 * j + d] + e - 1) / kCell;
            }
            for (int cx = c_lo[0]; cx <= c_hi[0]; cx++)
                for (int cy = c_lo[1]; cy <= c_hi[1]; cy++)
                    for (int cz = c_lo[2]; cz <= c_hi[2]; cz++)
                    {
                        const int cg[3] = { cx, cy, cz };
                        int rec[10];
                        rec[0] = j;
                        for (int d = 0; d < 3; d++)
                        {
                            const int o = s->origin[3 * j + d];
                            rec[1 + d] = kCell * cg[d] - o;
                            rec[4 + d] = std::max(0, o - kCell * cg[d]);
                            rec[7 + d] = std::min(kCell, o + e - kCell * cg[d]);
                        }
                        const size_t ci = ((size_t)(cx % C0) * C1 + (cy % C1)) * C2 + (cz % C2);
                        cl[ci].insert(cl[ci].end(), rec, rec + 10);
                    }
        }
        std::vector<int> cptr(1, 0), cent;
        std::vector<int64_t> cslot;
        std::vector<std::vector<int>> cells_of;  // box ids per kept cell (ascending)
        int64_t slot = 0;
        int kmax     = 0;
        for (auto& c : cl)
        {
            if (c.empty()) continue;
            const int k = (int)(c.size() / 10);  // boxes were appended in ascending order
            kmax        = std::max(kmax, k);
            cent.insert(cent.end(), c.begin(), c.end());
            cptr.push_back((int)(cent.size() / 10));
            cslot.push_back(slot);
            slot += (int64_t)k * k;
        }
        if (kmax <= kCellK)
        {
            // map (i, j) -> pair index
            std::vector<std::vector<std::pair<int, int>>> pidx(s->nbox);
            for (int64_t q = 0; q < s->npairs; q++)
                pidx[pairs[11 * q]].push_back({ pairs[11 * q + 1], (int)q });
            std::vector<int> pcount(s->npairs + 1, 0);
            std::vector<int64_t> pslot;
            std::vector<std::vector<int64_t>> lists(s->npairs);
            for (size_t c = 0; c + 1 < cptr.size(); c++)
            {
                const int k = cptr[c + 1] - cptr[c];
                for (int a = 0; a < k; a++)
                    for (int b = a; b < k; b++)
                    {
                        const int i = cent[10 * (cptr[c] + a)], j = cent[10 * (cptr[c] + b)];
                        auto& v   = pidx[i];
                        auto it   = std::lower_bound(v.begin(), v.end(), std::make_pair(j, -1));
                        if (it == v.end() || it->first != j) continue;  // cannot happen: covering the same cell
                        lists[it->second].push_back(cslot[c] + (int64_t)a * k + b);
                        lists[it->second].push_back(cslot[c] + (int64_t)b * k + a);
                    }
            }
            std::vector<int> pptr(s->npairs + 1, 0);
            for (int64_t q = 0; q < s->npairs; q++)
            {
                pptr[q + 1] = pptr[q] + (int)(lists[q].size() / 2);
                pslot.insert(pslot.end(), lists[q].begin(), lists[q].end());
            }
            s->ncell = (int)cslot.size();
            s->kmax  = kmax;
            s->nslot = slot;
            bool ok  = hipMalloc(&s->d_cell_ptr, sizeof(int) * cptr.size()) == hipSuccess
                      && hipMalloc(&s->d_cell_ent, sizeof(int) * std::max<size_t>(cent.size(), 1)) == hipSuccess
                      && hipMalloc(&s->d_cell_slot, sizeof(int64_t) * std::max<size_t>(cslot.size(), 1)) == hipSuccess
                      && hipMalloc(&s->d_pslot_ptr, sizeof(int) * pptr.size()) == hipSuccess
                      && hipMalloc(&s->d_pslot, sizeof(int64_t) * std::max<size_t>(pslot.size(), 1)) == hipSuccess;
            if (!ok) return fail("hipMalloc cell tables");
            hipMemcpy(s->d_cell_ptr, cptr.data(), sizeof(int) * cptr.size(), hipMemcpyHostToDevice);
            if (!cent.empty()) hipMemcpy(s->d_cell_ent, cent.data(), sizeof(int) * cent.size(), hipMemcpyHostToDevice);
            {
                // the entries with the box offsets and edges resolved
                std::vector<CellEnt> ce(cent.size() / 10);
                for (size_t q = 0; q < ce.size(); q++)
                {
                    const int* E = &cent[10 * q];
                    CellEnt c;
                    c.base = s->off[0][E[0]];
                    c.e    = s->edge[E[0]];
                    c.b0   = E[1];
                    c.b1   = E[2];
                    c.b2   = E[3];
                    c.lo0  = E[4];
                    c.lo1  = E[5];
                    c.lo2  = E[6];
                    c.hi0  = E[7];
                    c.hi1  = E[8];
                    c.hi2  = E[9];
                    c.box  = E[0];
                    ce[q]  = c;
                }
                if (hipMalloc(&s->d_cell_ce, sizeof(CellEnt) * std::max<size_t>(ce.size(), 1)) != hipSuccess)
                    return fail("hipMalloc cell entries");
                if (!ce.empty())
                    hipMemcpy(s->d_cell_ce, ce.data(), sizeof(CellEnt) * ce.size(), hipMemcpyHostToDevice);
            }
            if (!cslot.empty())
                hipMemcpy(s->d_cell_slot, cslot.data(), sizeof(int64_t) * cslot.size(), hipMemcpyHostToDevice);
            hipMemcpy(s->d_pslot_ptr, pptr.data(), sizeof(int) * pptr.size(), hipMemcpyHostToDevice);
            if (!pslot.empty())
                hipMemcpy(s->d_pslot, pslot.data(), sizeof(int64_t) * pslot.size(), hipMemcpyHostToDevice);
            {
                std::vector<std::vector<int>> be(s->nbox);
                const int nent = (int)(cent.size() / 10);
                for (int q = 0; q < nent; q++)
                    be[cent[10 * q]].push_back(q);
                std::vector<int> bptr(s->nbox + 1, 0), bl;
                for (int j = 0; j < s->nbox; j++)
                {
                    bptr[j + 1] = bptr[j] + (int)be[j].size();
                    bl.insert(bl.end(), be[j].begin(), be[j].end());
                }
                if (hipMalloc(&s->d_bent_ptr, sizeof(int) * bptr.size()) != hipSuccess
                    || hipMalloc(&s->d_bent, sizeof(int) * std::max<size_t>(bl.size(), 1)) != hipSuccess)
                    return fail("hipMalloc box entry lists");
                hipMemcpy(s->d_bent_ptr, bptr.data(), sizeof(int) * bptr.size(), hipMemcpyHostToDevice);
                if (!bl.empty()) hipMemcpy(s->d_bent, bl.data(), sizeof(int) * bl.size(), hipMemcpyHostToDevice);
                s->nent = nent;
            }
            s->use_cells = true;
            s->pair_z    = true;
            for (int j = 0; j < s->nbox; j++)
                s->pair_z = s->pair_z && s->edge[j] % 2 == 0 && s->origin[3 * j + 2] % 2 == 0;
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail("box table upload");
    *out = s;
    return MGX_OK;
}

int mgx_boxset_autotune(mgx_ctx* ctx, mgx_boxset* bs, int* choice, double* ms)
{
    if (!ctx) return MGX_EINVAL;
    MGX_REQUIRE(ctx, bs, "null box set");
    int rc = box_autotune(ctx, bs, ms);
    if (rc) return rc;
    if (choice)
        for (int t = 0; t < MGX_TUNE_N; t++)
            choice[t] = bs->tune[t];
    return MGX_OK;
}

int mgx_boxset_set_tune(mgx_boxset* bs, const int* choice)
{
    if (!bs || !choice) return MGX_EINVAL;
    for (int t = 0; t < MGX_TUNE_N; t++)
        if (choice[t] != 1 && choice[t] != 2) return MGX_EINVAL;
    for (int t = 0; t < MGX_TUNE_N; t++)
        bs->tune[t] = choice[t];
    bs->tuned = true;
    return MGX_OK;
}

int mgx_boxset_get_tune(const mgx_boxset* bs, int* choice, int* tuned)
{
    if (!bs) return MGX_EINVAL;
    if (choice)
        for (int t = 0; t < MGX_TUNE_N; t++)
            choice[t] = bs->tune[t];
    if (tuned) *tuned = bs->tuned ? 1 : 0;
    return MGX_OK;
}

int mgx_boxset_destroy(mgx_boxset* s)
{
    if (!s) return MGX_OK;
    for (auto p : s->d_off)
        if (p) hipFree(p);
    for (auto p : s->d_tile)
        if (p) hipFree(p);
    if (s->d_origin) hipFree(s->d_origin);
    if (s->d_edge) hipFree(s->d_edge);
    if (s->d_pairs) hipFree(s->d_pairs);
    if (s->d_nbr_ptr) hipFree(s->d_nbr_ptr);
    if (s->d_nbr) hipFree(s->d_nbr);
    if (s->d_cell_ptr) hipFree(s->d_cell_ptr);
    if (s->d_cell_ent) hipFree(s->d_cell_ent);
    if (s->d_cell_ce) hipFree(s->d_cell_ce);
    if (s->d_cell_slot) hipFree(s->d_cell_slot);
    if (s->d_pslot_ptr) hipFree(s->d_pslot_ptr);
    if (s->d_pslot) hipFree(s->d_pslot);
    if (s->d_ilist_ptr) hipFree(s->d_ilist_ptr);
    if (s->d_ilist) hipFree(s->d_ilist);
    if (s->d_pair_ient) hipFree(s->d_pair_ient);
    for (int* p : { s->d_sp_ptr, s->d_sp_col, s->d_sp_k, s->d_sp_e, s->d_bent_ptr, s->d_bent })
        if (p) hipFree(p);
    delete s;
    return MGX_OK;
}

int mgx_boxset_size(const mgx_boxset* bs, int level, int64_t* n)
{
    if (!bs || !n || level < 0 || level >= bs->levels) return MGX_EINVAL;
    *n = bs->off[level][bs->nbox];
    return MGX_OK;
}

int mgx_boxset_offset(const mgx_boxset* bs, int level, int j, int64_t* o)
{
    if (!bs || !o || level < 0 || level >= bs->levels || j < 0 || j > bs->nbox) return MGX_EINVAL;
    *o = bs->off[level][j];
    return MGX_OK;
}

int mgx_boxset_npairs(const mgx_boxset* bs, int64_t* n)
{
    if (!bs || !n) return MGX_EINVAL;
    *n = bs->npairs;
    return MGX_OK;
}

#define BOX_ARGS_OK(ctx, bs, level)                                                           \
    MGX_REQUIRE(ctx, bs != nullptr, "null box set");                                         \
    MGX_REQUIRE(ctx, (level) >= 0 && (level) < bs->levels, "level outside the box set")

int mgx_box_lap(mgx_ctx* ctx, const mgx_boxset* bs, int level, int type, int dtype, const void* in, void* out)
{
    BOX_ARGS_OK(ctx, bs, level);
    MGX_REQUIRE(ctx, in && out, "null data");
    int rc = dtype == MGX_F32 ? box_lap_t(ctx, bs, level, type, (const float*)in, (float*)out)
                              : box_lap_t(ctx, bs, level, type, (const double*)in, (double*)out);
    return rc ? rc : sync(ctx);
}

int mgx_box_hloc(mgx_ctx* ctx, const mgx_boxset* bs, int dtype, const void* psi, const double* vtot, void* hpsi)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, psi && vtot && hpsi, "null data");
    int rc = box_hloc(ctx, bs, dtype, psi, vtot, hpsi);
    return rc ? rc : sync(ctx);
}

int mgx_box_jacobi(mgx_ctx* ctx, const mgx_boxset* bs, int level, int type, int dtype, double jf, const void* f,
    void* x, void* w)
{
    BOX_ARGS_OK(ctx, bs, level);
    MGX_REQUIRE(ctx, f && x && w, "null data");
    const int64_t n = bs->off[level][bs->nbox];
    void* xo        = workspace(ctx, 5, n * esize(dtype));
    if (!xo) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    int rc = dtype == MGX_F32
                 ? box_jacobi_t(ctx, bs, level, type, jf, (const float*)x, (const float*)f, (float*)xo, (float*)w)
                 : box_jacobi_t(ctx, bs, level, type, jf, (const double*)x, (const double*)f, (double*)xo, (double*)w);
    if (rc) return rc;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(x, xo, n * esize(dtype), hipMemcpyDeviceToDevice, ctx->stream));
    return sync(ctx);
}

int mgx_box_restrict(mgx_ctx* ctx, const mgx_boxset* bs, int level, int dtype, const void* fine, void* coarse)
{
    BOX_ARGS_OK(ctx, bs, level + 1);
    MGX_REQUIRE(ctx, fine && coarse, "null data");
    int rc = dtype == MGX_F32 ? box_restrict_t(ctx, bs, level, (const float*)fine, (float*)coarse)
                              : box_restrict_t(ctx, bs, level, (const double*)fine, (double*)coarse);
    return rc ? rc : sync(ctx);
}

int mgx_box_prolong(mgx_ctx* ctx, const mgx_boxset* bs, int level, int dtype, const void* coarse, void* fine)
{
    BOX_ARGS_OK(ctx, bs, level + 1);
    MGX_REQUIRE(ctx, fine && coarse, "null data");
    int rc = dtype == MGX_F32 ? box_prolong_t<float, false>(ctx, bs, level, (const float*)coarse, (float*)fine)
                              : box_prolong_t<double, false>(ctx, bs, level, (const double*)coarse, (double*)fine);
    return rc ? rc : sync(ctx);
}

int mgx_box_precond_mg(mgx_ctx* ctx, const mgx_boxset* bs, int type, int max_levels, double gamma, int prec,
    const double* res, double* out)
{
    BOX_ARGS_OK(ctx, bs, max_levels);
    MGX_REQUIRE(ctx, res && out, "null data");
    MGX_REQUIRE(ctx, prec == MGX_F32 || prec == MGX_F64 || prec == MGX_F16, "bad precision");
    int rc;
    if (prec == MGX_F16)
    {
        const int64_t n = bs->off[0][bs->nbox];
        float* x32      = (float*)workspace(ctx, 5, n * sizeof(float));
        if (!x32) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
        if ((rc = box_precond_h(ctx, bs, type, max_levels, gamma, res, nullptr, x32, nullptr, nullptr))) return rc;
        hipLaunchKernelGGL(k_to_f64<float>, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, (const float*)x32, out, n);
        MGX_CHECK_LAUNCH(ctx);
        return sync(ctx);
    }
    rc = box_precond(ctx, bs, type, max_levels, gamma, prec, res, out, nullptr, 0., false);
    return rc ? rc : sync(ctx);
}

int mgx_box_gram_h(mgx_ctx* ctx, const mgx_boxset* bs, int prec, const double* psi, const double* hpsi, double* S,
    double* H)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, psi && hpsi && S && H, "null data");
    MGX_REQUIRE(ctx, prec == MGX_PREC_F64 || prec == MGX_PREC_SPLIT, "bad precision mode");
    int rc = box_gram_h(ctx, bs, prec, psi, hpsi, S, H, nullptr, nullptr);
    return rc ? rc : sync(ctx);
}

int mgx_box_residual(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi, const double* theta,
    double* res)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, psi && hpsi && theta && res, "null data");
    int rc = box_residual(ctx, bs, psi, hpsi, theta, res, nullptr);
    return rc ? rc : sync(ctx);
}

int mgx_box_hloc_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* vtot, double* hpsi)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, psi && vtot && hpsi, "null data");
    int rc = box_hloc_mehr(ctx, bs, psi, vtot, hpsi);
    return rc ? rc : sync(ctx);
}

int mgx_box_gram_hb_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi,
    const double* vtot, double* S, double* H, double* B)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, psi && hpsi && vtot && S && H && B, "null data");
    const int64_t np = bs->off[0][bs->nbox];
    double* bpsi     = (double*)workspace(ctx, 5, np * sizeof(double));
    if (!bpsi) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    int rc = box_bpsi_mehr(ctx, bs, psi, bpsi);
    if (!rc) rc = box_gram_hb_mehr(ctx, bs, psi, hpsi, bpsi, vtot, S, H, B);
    return rc ? rc : sync(ctx);
}

int mgx_box_residual_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi,
    const double* theta, double* res)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, psi && hpsi && theta && res, "null data");
    const int64_t np = bs->off[0][bs->nbox];
    double* bpsi     = (double*)workspace(ctx, 5, np * sizeof(double));
    if (!bpsi) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    int rc = box_bpsi_mehr(ctx, bs, psi, bpsi);
    if (!rc) rc = box_residual_mehr(ctx, bs, psi, bpsi, hpsi, theta, res);
    return rc ? rc : sync(ctx);
}

int mgx_box_density(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* dm, int n, double nel,
    double* rho, double* charge)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, psi && dm && rho, "null data");
    int rc = box_density(ctx, bs, psi, dm, n, nel, rho, charge);
    return rc ? rc : sync(ctx);
}

int mgx_box_residual_norm(mgx_ctx* ctx, const mgx_boxset* bs, const double* res, const double* Sinv, double* norm)
{
    BOX_ARGS_OK(ctx, bs, 0);
    MGX_REQUIRE(ctx, res && Sinv && norm, "null data");
    double* d = (double*)workspace(ctx, 6, (bs->nbox + 8) * sizeof(double));
    if (!d) return set_err(ctx, MGX_ENOMEM, "workspace allocation failed");
    int rc = box_sumsq(ctx, bs, res, d);
    if (!rc) rc = box_norm(ctx, bs, d, Sinv, d + bs->nbox);
    if (rc) return rc;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(norm, d + bs->nbox, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return sync(ctx);
}

}  // extern "C"
