// common.hpp -- context, error handling and launch helpers shared by the
// MI355X (gfx950) kernels of the orbital-iteration engine.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "mgx.h"

// hipEvent pairs around kernel groups, recorded on the context stream;
// collected (elapsed ms accumulated per group) after a synchronisation.
// Each pair also carries the algorithmic work of what it brackets (HBM bytes
// or flops, the unit fixed per group), so that rates come out per group.
constexpr int kProfGroups = 32;
struct MgxProf
{
    std::vector<hipEvent_t> pool;
    std::vector<int> group;     // group of each recorded pair
    std::vector<double> work;   // algorithmic work of each recorded pair
    size_t used = 0;            // events handed out
    double ms[kProfGroups]   = {};
    int64_t cnt[kProfGroups] = {};
    double wsum[kProfGroups] = {};
};

struct mgx_ctx
{
    int device       = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    std::string err;
    // grow-only device workspaces; a call owns them for its duration (the
    // context is not re-entrant), slot = independent buffer
    static constexpr int kSlots = 10;
    void* ws[kSlots]     = {};
    size_t ws_sz[kSlots] = {};
    uint64_t ws_gen      = 0;  // bumped whenever a workspace moves (captured graphs hold the old address)
    double* host_pinned  = nullptr;  // small pinned host buffer for scalars
    MgxProf* prof        = nullptr;  // non-null while a profiled solver runs
};

namespace mgx
{

int set_err(mgx_ctx* ctx, int code, const std::string& msg);

// device workspace of at least `bytes` in slot `slot` (256-byte aligned)
void* workspace(mgx_ctx* ctx, int slot, size_t bytes);

// device address of this GPU's no-allocation scratch array (dense_ops.hip);
// called once per device at context creation, with the device current
int scratch_init(mgx_ctx* ctx, int device);

inline size_t esize(int dtype) { return dtype == MGX_F32 ? sizeof(float) : sizeof(double); }

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n, int per_block = kBlock, int64_t cap = 1 << 20)
{
    int64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// blocks are dispatched round-robin over the 8 XCDs of the MI355X: remap so
// that consecutive logical blocks share an XCD (and its L2).
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nb)
{
    const int64_t per = (nb + 7) / 8;
    const int64_t xcd = bid % 8;
    const int64_t k   = bid / 8;
    int64_t logical   = xcd * per + k;
    return logical;
}

struct GGeom
{
    int d0, d1, d2, ng;
    int64_t incx, incy, sizeg, size;
};

inline GGeom ggeom(const mgx_grid_t& g)
{
    GGeom r;
    r.d0    = g.dim[0];
    r.d1    = g.dim[1];
    r.d2    = g.dim[2];
    r.ng    = g.ghosts;
    r.incy  = g.dim[2] + 2 * g.ghosts;
    r.incx  = (int64_t)(g.dim[1] + 2 * g.ghosts) * r.incy;
    r.sizeg = (int64_t)(g.dim[0] + 2 * g.ghosts) * r.incx;
    r.size  = (int64_t)g.dim[0] * g.dim[1] * g.dim[2];
    return r;
}

// finite-difference coefficients, computed on the host exactly as the
// reference does (FDkernels.cc:92-104, 457-463; Laph2: FDkernels.cc:30-43)
struct Coef4
{
    double c0, c1x, c2x, c1y, c2y, c1z, c2z;
};
struct Coef2
{
    double c0, c1x, c1y, c1z;
};
struct CoefM
{
    double c0, cx, cy, cz, cxy, cyz, cxz;
};

inline Coef4 coef4(const double h[3])
{
    const double inv12 = 1. / 12.;
    double ih[3];
    for (int i = 0; i < 3; i++)
        ih[i] = 1. / (h[i] * h[i]);
    Coef4 c;
    const double cc0 = inv12 * ih[0];
    c.c1x            = -16. * cc0;
    c.c2x            = 1. * cc0;
    const double cc1 = inv12 * ih[1];
    c.c1y            = -16. * cc1;
    c.c2y            = 1. * cc1;
    const double cc2 = inv12 * ih[2];
    c.c1z            = -16. * cc2;
    c.c2z            = 1. * cc2;
    c.c0             = -2. * (c.c1x + c.c2x + c.c1y + c.c2y + c.c1z + c.c2z);
    return c;
}

inline Coef2 coef2(const double h[3])
{
    Coef2 c;
    c.c1x = -(1. / (h[0] * h[0]));
    c.c1y = -(1. / (h[1] * h[1]));
    c.c1z = -(1. / (h[2] * h[2]));
    c.c0  = -2. * (c.c1x + c.c1y + c.c1z);
    return c;
}

inline CoefM coefm(const double h[3])
{
    const double inv12 = 1. / 12.;
    double ih[3];
    for (int i = 0; i < 3; i++)
        ih[i] = 1. / (h[i] * h[i]);
    CoefM c;
    c.c0  = 16. * inv12 * (ih[0] + ih[1] + ih[2]);
    c.cx  = -10. * inv12 * ih[0] + 0.125 * c.c0;
    c.cy  = -10. * inv12 * ih[1] + 0.125 * c.c0;
    c.cz  = -10. * inv12 * ih[2] + 0.125 * c.c0;
    c.cxy = -inv12 * (ih[0] + ih[1]);
    c.cyz = -inv12 * (ih[2] + ih[1]);
    c.cxz = -inv12 * (ih[0] + ih[2]);
    return c;
}

// Laph*::diagEl and jacobiFactor: src/pb/Laph4M.h:29-31,80; Laph2.h:32-33,95;
// Laph4.h:35-38,106.
inline double diag_el(int type, const double h[3])
{
    const double s = 1. / (h[0] * h[0]) + 1. / (h[1] * h[1]) + 1. / (h[2] * h[2]);
    if (type == MGX_LAP_MEHRSTELLEN) return (4. / 3.) * s;
    if (type == MGX_LAP_2ND) return 2. * s;
    if (type == MGX_LAP_4TH) return 2.5 * s;
    return 0.;
}

inline double jacobi_factor(int type, const double h[3])
{
    const double inv = 1. / diag_el(type, h);
    if (type == MGX_LAP_MEHRSTELLEN) return inv;
    if (type == MGX_LAP_2ND) return 6. * inv / 7.;
    if (type == MGX_LAP_4TH) return inv / 1.5;
    return 0.;
}

int sync(mgx_ctx* ctx);

// profiling tokens (no-ops when ctx->prof is null); work = the bracketed
// launches' algorithmic bytes or flops (the group's unit)
int prof_begin(mgx_ctx* ctx, int group, double work = 0.);
void prof_end(mgx_ctx* ctx, int token);
// change the group and work of a recorded token (once a launch knows what it did)
void prof_retag(mgx_ctx* ctx, int token, int group, double work);
void prof_collect(mgx_ctx* ctx);

// Column tiles of a box set's pair-sparse matrices (S and H have a nonzero
// (i, j) only for interacting pairs): tile t covers kSpTile columns (boxes
// grouped by space), ptr[t]..ptr[t+1] index the union of their rows k, and
// e[kSpTile*u + q] is the pair whose value sits at (k_u, col_q), or -1.
constexpr int kSpTile = 4;  // measured at C5: 4 columns per tile 0.90, 2: 0.92, 8: 0.97, 16: 1.08 ms of dense phase per step
struct PairTiles
{
    int ntile      = 0;
    const int* ptr = nullptr;
    const int* col = nullptr;  // kSpTile per tile, -1 padded
    const int* k   = nullptr;
    const int* e   = nullptr;
};

// Y = X A for dense symmetric-storage X (n x n, column-major) and the
// pair-sparse symmetric A given by its per-pair values
int spmm_pairs(mgx_ctx* ctx, int n, const double* X, const double* val, const PairTiles& pt, double* Y);

// grid-family internals reused by the solver and tests
int halo(mgx_ctx* ctx, int dtype, void* data, int nfunc, const mgx_grid_t& g);

}  // namespace mgx

#define MGX_CHECK_HIP(ctx, call)                                                               \
    do                                                                                         \
    {                                                                                          \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return mgx::set_err((ctx), MGX_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MGX_CHECK_LAUNCH(ctx) MGX_CHECK_HIP(ctx, hipGetLastError())

#define MGX_REQUIRE(ctx, cond, msg)                                                            \
    do                                                                                         \
    {                                                                                          \
        if (!(cond)) return mgx::set_err((ctx), MGX_EINVAL, (msg));                            \
    } while (0)
