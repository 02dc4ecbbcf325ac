/*
 * mgx.h -- C ABI of the MI355X engine for MGmol's orbital-iteration hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (MGX_OK or a negative MGX_E* code; mgx_last_error() gives the
 * text) and never throws.  Field and matrix pointers are DEVICE pointers
 * (HIP, gfx950) unless a parameter says "host".  Calls are ordered on the
 * context's stream and return after the work completes (synchronous), so a
 * caller may read results immediately.  A context is not re-entrant: use one
 * per host thread.
 *
 * Layout (src/pb/Grid.cc:80-82): a grid function has x slowest, z fastest;
 * a ghosted block stores nfunc functions of sizeg = prod(dim+2*ghosts)
 * points each; a packed block stores nfunc functions of numpt = prod(dim)
 * points each with leading dimension ld (LocGridOrbitals / BlockVector
 * storage, src/BlockVector.h:59-67).  Small matrices are column-major FP64.
 *
 * Each declaration names the reference interface it replaces (file:line).
 */
#ifndef MGX_H
#define MGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGX_OK 0
#define MGX_EINVAL -1   /* bad argument (reference: assert / exit(2)) */
#define MGX_EHIP -2     /* HIP runtime or library failure */
#define MGX_ENOTSPD -3  /* Cholesky failed: S not SPD (reference: mmpi.abort) */
#define MGX_ENUMERIC -4 /* NaN/Inf detected, or a density charge off the electron count */
#define MGX_ENOMEM -5   /* device allocation failed */

#define MGX_F32 0
#define MGX_F64 1
/* preconditioner storage only (mgx_box_precond_mg, mgx_solver_cfg_t.precond_dtype):
   the V-cycle in FP16 with the residual scaled per box, its correction in FP32 */
#define MGX_F16 2

/* lap_type values of LapFactory<T>::createLap, src/LapFactory.h:26-56 */
#define MGX_LAP_MEHRSTELLEN 0
#define MGX_LAP_2ND 1
#define MGX_LAP_4TH 2

/* precision of the S / H contractions */
#define MGX_PREC_F64 0   /* FP64 MFMA, FP64 accumulation */
#define MGX_PREC_SPLIT 1 /* FP32 MFMA on hi/lo split, FP32 chunk sums, FP64 total */

typedef struct mgx_ctx mgx_ctx;
typedef struct mgx_boxset mgx_boxset;
typedef struct mgx_solver mgx_solver;

/* pb::Grid, src/pb/Grid.h:24-121 (one task owns the whole grid) */
typedef struct
{
    int dim[3];  /* interior points */
    int ghosts;  /* ghost width */
    double h[3]; /* spacing */
    int bc[3];   /* 1 = periodic; 0 = Dirichlet as GridFunc::setBoundaryValues
                  * (src/pb/GridFunc.cc:2189-2340) sets it: the ghosts and the
                  * index-0 layer (the boundary plane) hold zero, a halo
                  * exchange zeroes that layer in place, and the kernels that
                  * wrap in place read it as zero; interior norms leave it out */
} mgx_grid_t;

/* pb::GridFuncVector<T> storage, src/pb/GridFuncVector.h:30-493 */
typedef struct
{
    void* data; /* device, nfunc*sizeg elements of dtype */
    int dtype;  /* MGX_F32 | MGX_F64 */
    int nfunc;
    mgx_grid_t grid;
} mgx_block_t;

/* per-orbital cubic boxes on a periodic grid (0/1 cube masks of
 * Map2Masks, src/Map2Masks.cc:26-58).  origin/edge are HOST arrays. */
typedef struct
{
    int nbox;
    int gdim[3];       /* global grid at level 0 */
    double h[3];       /* spacing at level 0 */
    const int* origin; /* host, 3*nbox, multiples of 2^(levels-1) */
    const int* edge;   /* host, nbox, multiples of 2^(levels-1), 2*edge+4 <= gdim */
} mgx_boxes_t;

/* ------------------------------------------------------------ context ---- */
/* Device selection and stream: replaces MagmaSingleton (src/magma_singleton.cc) */
int mgx_ctx_create(int device, mgx_ctx** ctx);
int mgx_ctx_destroy(mgx_ctx* ctx);
const char* mgx_last_error(const mgx_ctx* ctx);
/* hipStream_t of the context (for callers that interleave their own work) */
void* mgx_ctx_stream(mgx_ctx* ctx);

/* device memory: MemorySpace::Memory<T, Device>::allocate/free and
 * MemorySpace::copy_to_dev / copy_to_host, src/tools/memory_space.h */
int mgx_malloc(mgx_ctx* ctx, size_t bytes, void** ptr);
int mgx_free(mgx_ctx* ctx, void* ptr);
int mgx_memcpy_htod(mgx_ctx* ctx, void* dst_device, const void* src_host, size_t bytes);
int mgx_memcpy_dtoh(mgx_ctx* ctx, void* dst_host, const void* src_device, size_t bytes);

/* -------------------------------------------- grid operators (ghosted) ---- */
/* GridFuncVector::trade_boundaries, src/pb/GridFuncVector.cc:1544-1622 */
int mgx_halo(mgx_ctx* ctx, mgx_block_t* b);
/* FDkernelDel2_4th_Mehr / _2nd / _4th via GridFuncVector::applyLap,
 * src/pb/GridFuncVector.cc:2370-2397; src/pb/FDkernels.cc:19-162,448-520.
 * Reads the ghosts of `in` as stored (trade first), writes the interior. */
int mgx_lap_apply(mgx_ctx* ctx, int lap_type, const mgx_block_t* in, mgx_block_t* out);
/* FDkernelRHS_4th_Mehr1, src/pb/FDkernels.cc:523-584 (out ghost width free) */
int mgx_rhs_apply(mgx_ctx* ctx, const mgx_block_t* in, mgx_block_t* out);
/* GridFuncVector::pointwiseProduct, src/pb/GridFuncVector.cc:90-136;
 * v: device FP64 ghosted single function */
int mgx_pointwise_product(mgx_ctx* ctx, const mgx_block_t* a, const double* v, mgx_block_t* out);
/* Synthetic input generator for the dense contraction workload (no
 * counterpart in the reference): separable orbitals psi_j(x,y,z) =
 * fx_j(x) fy_j(y) fz_j(z) on the z-planes [z0, z0 + nzs) of a periodic
 * gdim grid, and H psi_j = (-Delta_h4 + V) psi_j from the factors' 1D
 * 4th-order second differences lf (the stencil is a sum of 1D stencils).
 * Factors n x gdim[d] row-major; vtot gdim[0]*gdim[1]*gdim[2] (z fastest);
 * psi, hpsi column-major, ld = gdim[0]*gdim[1]*nzs (x slowest, z fastest). */
int mgx_synth_separable(mgx_ctx* ctx, int n, const int gdim[3], int z0, int nzs, const double* fx,
    const double* fy, const double* fz, const double* lfx, const double* lfy, const double* lfz,
    const double* vtot, double* psi, double* hpsi);
/* LinearAlgebraUtils::MPaxpy, src/linear_algebra/mputils.cc:221-244:
 * y += (T)(alpha*(double)x), n elements of dtype */
int mgx_axpy(mgx_ctx* ctx, int dtype, int64_t n, double alpha, const void* x, void* y);
/* No counterpart in the reference: the HBM ceiling for the roofline fractions.
 * STREAM triad a = b + 0.5 c over n FP32 values (n % 4 == 0, 16-byte aligned
 * device arrays), 16-byte lanes with 2, 4 or 8 loads of each operand in
 * flight per lane, reps launches each timed with hipEvents on the context
 * stream; *gbps = the best rate (12 n bytes per launch), *unroll = its depth. */
int mgx_stream_triad(mgx_ctx* ctx, int64_t n, float* a, const float* b, const float* c, int reps,
    double* gbps, int* unroll);
/* GridFuncVector::jacobi, src/pb/GridFuncVector.cc:2416-2425 (trades x):
 * w = L x; w -= f; x += -jacobi_factor * w */
int mgx_jacobi(mgx_ctx* ctx, int lap_type, double jacobi_factor, const mgx_block_t* f,
    mgx_block_t* x, mgx_block_t* w);
/* GridFuncVector::restrict3D -> MGkernelRestrict3D, src/pb/MGkernels.cc:204-283
 * (fine ghosts must be traded) */
int mgx_restrict(mgx_ctx* ctx, const mgx_block_t* fine, mgx_block_t* coarse);
/* GridFuncVector::extend3D -> MGkernelExtend3D, src/pb/MGkernels.cc:29-201
 * (coarse ghosts must be traded) */
int mgx_prolong(mgx_ctx* ctx, const mgx_block_t* coarse, mgx_block_t* fine);
/* GridFuncVector::app_mask -> Map2Masks::apply, src/pb/GridFuncVector.cc:2428-2438;
 * mask: device uint8, nfunc*numpt (0 zeroes the interior point) */
int mgx_mask(mgx_ctx* ctx, mgx_block_t* u, const uint8_t* mask);
/* Preconditioning<T>::mg, src/Preconditioning.cc:155-216: V-cycle on v for
 * right-hand side f (same dtype); masks: NULL or host array of max_levels+1
 * device pointers (each NULL or nfunc*numpt_level bytes) */
int mgx_mg(mgx_ctx* ctx, int lap_type, int max_levels, mgx_block_t* v, const mgx_block_t* f,
    const uint8_t* const* masks);

/* Hartree potential by multigrid: pb::Mgm (src/pb/Mgm.h:12-111) with pb::Vcycle
 * (src/pb/Vcycle.h:28-249), as pb::SolverLap::solve runs them
 * (src/pb/SolverLap.cc:49-63) for Hartree::solve (src/Hartree.cc:27-100; the
 * caller passes rho already multiplied by 4 pi).  rhs = A.rhs(rho) (the
 * Mehrstellen B operator for lap_type 0, rho itself otherwise); up to max_sweeps
 * times: stop when |rhs - A vh| / |rhs| < tol, else one V-cycle on the residual
 * (nu1 / nu2 Jacobi sweeps, coarse levels while every dimension is even and at
 * least twice the ghost width, at most max_nlevels coarsenings, 4th order coarse
 * levels in 2nd order) added to vh; then, on a fully periodic grid,
 * vh -= mean(vh) (GridFunc::average0, src/pb/SolverLap.cc:60), and its ghosts
 * are refreshed.  rho, vh: one FP64 function each, same grid, every direction
 * periodic (bc 1) or Dirichlet (bc 0, zero boundary values; the multipole
 * values of bc 2, src/Hartree.cc:50-76, are not built); vh is the initial
 * guess and the result.  The reference's
 * defaults are nu1 = nu2 = 2, max_sweeps = 10, tol = 1e-16, max_nlevels = 10. */
typedef struct
{
    int converged;
    int nb_sweeps;
    double final_residual;          /* pb::norm of rhs - A vh */
    double final_relative_residual; /* final_residual / |rhs| */
    double residual_reduction;      /* final_residual / initial residual */
} mgx_poisson_stats_t;
int mgx_poisson_mg(mgx_ctx* ctx, int lap_type, const mgx_block_t* rho, mgx_block_t* vh,
    int nu1, int nu2, int max_sweeps, double tol, int max_nlevels, mgx_poisson_stats_t* stats);
/* SolverPB<PBh4M / PBh4 / PBh2<double>, double>::solve(phi, rhs, rhod, vks),
 * src/pb/SolverPB.cc:56-84 (PBdiel::solve, src/PBdiel.cc:30-68, with the operator
 * PoissonSolverFactory::createDiel picks, src/PoissonSolverFactory.h:151-190):
 * the Poisson-Boltzmann potential in a
 * dielectric continuum eps(rhod) (DielFunc::epsilon_rho, src/pb/DielFunc.cc:38-43:
 * 1 + (e0-1)/2 ((1-r)/(1+r) + 1), r = (rhod/rho0)^(2 drho0)), as pb::Mgm solves it
 * with PBh4M (src/pb/PBh4M.h, src/pb/PBh4M.cc): vh *= sqrt(eps), rhs = B(rho /
 * sqrt(eps)), cycles of -Lap_M u + B(pp u), pp = Lap_4 sqrt(eps) / sqrt(eps), with
 * the PBh4M Jacobi smoother (x -= 0.75 / (c0 + pp) (A x - f), c0 = 30/12 sum 1/h^2)
 * and PBh4M on every coarse level (pp restricted), vh /= sqrt(eps) at the end (no
 * mean removal); then, if vks is not null, PBh4M::get_vepsilon: vks = d eps/d rho
 * |grad vh|^2 (-1/(8 pi)), the gradient of 4th order.  That is lap_type
 * MGX_LAP_MEHRSTELLEN.  MGX_LAP_4TH (PBh4, src/pb/PBh4.cc) and MGX_LAP_2ND (PBh2,
 * src/pb/PBh2.cc) take the gradient form on the untransformed potential: A x =
 * eps (-Lap x) - grad x . grad eps (first differences del1_4th / del1_2nd),
 * rhs = rho, the Jacobi smoother x -= (A x - f) / (c0 eps) with c0 = 1.5 (30/12)
 * sum 1/h^2, PBh2 on every coarse level (USE_LOWER_ORDER) on the restricted eps,
 * and get_vepsilon's gradient of the operator's order.  rho, rhod, vh, vks: one
 * FP64 function each, same grid, 2 ghosts at least (1 for MGX_LAP_2ND), bc per
 * direction 1 or 0; the dielectric functions are periodic whatever bc says.  With
 * PBh4M the coarse levels keep pp and the functions on the same points (the
 * reference pairs pp laid out with 2 ghosts with 1-ghost coarse functions element
 * by element). */
int mgx_pb_mg(mgx_ctx* ctx, int lap_type, const mgx_block_t* rho, const mgx_block_t* rhod,
    mgx_block_t* vh, mgx_block_t* vks, double e0, double rho0, double drho0, int nu1, int nu2,
    int max_sweeps, double tol, int max_nlevels, mgx_poisson_stats_t* stats);

/* ------------------------------------ Hamiltonian / preconditioner (packed) */
/* Hamiltonian<T>::applyLocal(ncolors, phi, hphi), src/Hamiltonian.cc:86-159
 * (4th order: FDoper::del2_4th_withPot, src/pb/FDoper.cc:322-399; Mehrstellen:
 * -Lap_M psi + B(V psi)).  psi/hpsi packed (numpt per function, lds), halo
 * from grid->bc; vtot device FP64 packed numpt (Rydberg). */
int mgx_hloc_apply(mgx_ctx* ctx, int lap_type, const mgx_grid_t* grid, int dtype, const void* psi,
    int64_t ld, int nfunc, const double* vtot, void* hpsi, int64_t ld_out);
/* OrbitalsPreconditioning<T>::precond_mg, src/OrbitalsPreconditioning.cc:87-117:
 * out = mg(x0 = gamma*res, f = res) in prec_dtype (MGX_F32 = MGPRECONDTYPE of
 * the reference, src/global.h:33); res/out FP64 packed.  gamma > 0 as in the
 * reference; the box variant below also accepts gamma = 0 (start from 0). */
int mgx_precond_mg(mgx_ctx* ctx, int lap_type, int max_levels, double gamma, int prec_dtype,
    const mgx_grid_t* grid, const double* res, int64_t ld, int nfunc, const uint8_t* const* masks,
    double* out, int64_t ld_out);
/* OrbitalsPreconditioning<T>::setGamma, src/OrbitalsPreconditioning.cc:120-145
 * (host computation): gamma = 1/(2/(invDiag(Laph4)*4^(mg_levels+1)) + |vmax - small_eig|) */
double mgx_gamma(const double h[3], int mg_levels, double vmax, double small_eig);

/* ------------------------------------------------------ dense FP64 layer */
/* LocGridOrbitals::getLocalOverlap -> LocalMatrices::syrk -> MPsyrk,
 * src/LocGridOrbitals.cc:1504-1530, src/local_matrices/LocalMatrices.cc:210-247:
 * ss[s] = vel * psi_s^T psi_s (full symmetric, n x n col-major) for slabs
 * s < subdivx of loc_numpt rows; psi column-major (ld). */
int mgx_local_gram(mgx_ctx* ctx, int prec, const double* psi, int64_t ld, int64_t loc_numpt, int n,
    int subdivx, double vel, double* ss);
/* No counterpart in the reference: mgx_local_gram for callers that cannot
 * allocate device memory (code under HIP graph capture, tight loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the chunk
 * partials live in the library's fixed 16 MiB device array (one per GPU,
 * kept until the process ends) instead of a workspace, so it calls no
 * hipMalloc/hipFree.  Limits: n < 256 and subdivx * n * n * 8 <= 16 MiB
 * (MGX_EINVAL beyond); calls that overlap on one GPU share the array and
 * race. */
int mgx_local_gram_noalloc(mgx_ctx* ctx, int prec, const double* psi, int64_t ld, int64_t loc_numpt, int n,
    int subdivx, double vel, double* ss);
/* LocGridOrbitals::computeLocalProduct -> LocalMatrices::gemm -> MPgemm('t','n'),
 * src/LocGridOrbitals.cc:1554-1604: ss[s](i,j) = vel * <a_i, b_j>_s */
int mgx_local_product(mgx_ctx* ctx, int prec, const double* a, const double* b, int64_t ld,
    int64_t loc_numpt, int n, int subdivx, double vel, double* ss);
/* LocalMatrices2DistMatrix::convert/accumulate, src/LocalMatrices2DistMatrix.cc:21-107:
 * global += slab entries (|x| > tol) at gid pairs; gids device int, subdivx*m */
int mgx_local_to_global(mgx_ctx* ctx, const double* ss, const int* gids, int subdivx, int m,
    int numst, double tol, double* global);
/* GramMatrix::updateLS + computeInverse, src/GramMatrix.cc:165-192, 82-93:
 * L = potrf('l', S) (lower, upper zeroed), Sinv = S^-1 (symmetric).
 * MGX_ENOTSPD if S is not positive definite. */
int mgx_chol_inv(mgx_ctx* ctx, int n, const double* S, double* L, double* Sinv);
/* ProjectedMatrices::updateTheta, src/ProjectedMatrices.h:218-225: theta = Sinv*H */
int mgx_theta(mgx_ctx* ctx, int n, const double* Sinv, const double* H, double* theta);
/* ProjectedMatrices::getEigSum, src/ProjectedMatrices.cc:550-569: tr(H Sinv) [Ry], host out */
int mgx_eigsum(mgx_ctx* ctx, int n, const double* H, const double* Sinv, double* trace);
/* ProjectedMatrices::solveGenEigenProblem, src/ProjectedMatrices.cc:277-295:
 * H z = lambda S z; eigvals ascending (device n), Z device n x n (may be NULL) */
int mgx_ritz(mgx_ctx* ctx, int n, const double* H, const double* S, double* eigvals, double* Z);
/* MGmol::computeResidualUsingHPhi (non-Mehrstellen), src/MGmol.cc:1227-1291:
 * res = psi*theta - hpsi (packed, ld, numpt rows, n functions) */
int mgx_residual(mgx_ctx* ctx, const double* psi, const double* hpsi, const double* theta,
    int64_t ld, int64_t numpt, int n, double* res);
/* Mehrstellen discretisation (FDtype = Mehrstellen) on packed orbitals (ld, numpt =
 * prod(grid->dim) rows, n functions, FP64; halo from grid->bc):
 * B-matrix: LocGridOrbitals::computeMatB, src/LocGridOrbitals.cc:1428-1486:
 *   B = vel * Psi^T (B Psi), B the Mehrstellen right-hand-side operator
 *   (FDkernelRHS_4th_Mehr1, src/pb/FDkernels.cc:523-584); n x n col-major */
int mgx_mehr_bmatrix(mgx_ctx* ctx, const mgx_grid_t* grid, const double* psi, int64_t ld, int n,
    double* B);
/* ProjectedMatricesMehrstellen::computeInvB / updateTheta / updateHB,
 * src/ProjectedMatricesMehrstellen.cc:38-64, .h:46-69, and getEigSum,
 * src/ProjectedMatrices.cc:550-569: invB = B^-1 (Cholesky; MGX_ENOTSPD),
 * theta = invB H, HB = (S theta + (S theta)^T) / 2, *eigsum (host, may be
 * NULL) = tr(HB Sinv) [Ry] */
int mgx_mehr_projected(mgx_ctx* ctx, int n, const double* S, const double* Sinv, const double* B,
    const double* H, double* invB, double* theta, double* HB, double* eigsum);
/* MGmol::computeResidualUsingHPhi with applyB, src/MGmol.cc:1227-1291:
 * res = (B psi) theta - hpsi */
int mgx_mehr_residual(mgx_ctx* ctx, const mgx_grid_t* grid, const double* psi, const double* hpsi,
    const double* theta, int64_t ld, int n, double* res);
/* MGmol::computeConstraintResidual with LocGridOrbitals::dotProductDiagonal,
 * src/MGmol.cc:1296-1332, src/LocGridOrbitals.cc:1606-1630:
 * 0.5*sqrt(sum_i Sinv_ii * vel*|r_i|^2) [Ha], host out */
int mgx_residual_norm(mgx_ctx* ctx, const double* res, int64_t ld, int64_t numpt, int n,
    double vel, const double* Sinv, double* norm);

/* --------------------------------------------------------- box family ---- */
int mgx_boxset_create(mgx_ctx* ctx, const mgx_boxes_t* boxes, int levels, mgx_boxset** bs);
/* Autotuner of the box set's fine-level kernels (no counterpart in the
 * reference).  Tunables (index into choice[MGX_TUNE_N]):
 *   MGX_TUNE_MARCH_F16 / _F32 / _F64: 4th-order Jacobi sweep of the marching
 *     kernel on FP16 / FP32 / FP64 values, MGX_TUNE_HLOC_F64: H psi (FP64):
 *     lane groups (16 bytes of z-points, 8 for FP16) per thread, 1 or 2; the
 *     block then has (e^2 / group size) / choice threads (64..1024);
 *   MGX_TUNE_GRAM_F64 / _SPLIT: waves per 8^3 cell of the S / H cell kernel
 *     with paired z-loads (FP64 / hi-lo split MFMA), 1 or 2;
 *   MGX_TUNE_RESTRICT: the V-cycle's restriction of the stale residual
 *     (MGkernelRestrict3D) inside the last pre-sweep of the marching kernel
 *     (1; w is then never written) or as its own kernel after it (2);
 *   MGX_TUNE_PROLONG: the coarse correction's prolongation
 *     (MGkernelExtend3D, v -= P(c)) inside the first post-sweep (1; v is
 *     then never written with it) or as its own kernel before it (2); both
 *     transfers timed on the FP32 fine sweeps, applied at every level and
 *     precision where the marching kernel runs the level.
 * mgx_boxset_autotune times every candidate on the box set's own layout
 * (scratch buffers of its size, 1 warm-up + 3 timed launches each, hipEvents)
 * and keeps the fastest; ms (may be NULL) receives MGX_TUNE_N x 2 times (ms
 * per launch, candidate 1 then 2; -1 where the kernel does not apply).  Any
 * choice gives the same results (the per-point arithmetic does not depend
 * on the lane mapping; only the Anderson dot sums of the last sweep are
 * ordered differently; the fused restriction sums its 27 FP64 terms in
 * another order: the same FP32 / FP16 values, FP64 ones to the last bit;
 * the fused prolongation is bit-identical).  A solver tunes its box set at creation unless it
 * was tuned or set before.  mgx_boxset_set_tune forces a choice; get reads it. */
#define MGX_TUNE_MARCH_F16 0
#define MGX_TUNE_MARCH_F32 1
#define MGX_TUNE_MARCH_F64 2
#define MGX_TUNE_HLOC_F64 3
#define MGX_TUNE_GRAM_F64 4
#define MGX_TUNE_GRAM_SPLIT 5
#define MGX_TUNE_RESTRICT 6
#define MGX_TUNE_PROLONG 7
#define MGX_TUNE_N 8
int mgx_boxset_autotune(mgx_ctx* ctx, mgx_boxset* bs, int* choice, double* ms);
int mgx_boxset_set_tune(mgx_boxset* bs, const int* choice);
int mgx_boxset_get_tune(const mgx_boxset* bs, int* choice, int* tuned);
int mgx_boxset_destroy(mgx_boxset* bs);
/* total points of all boxes at a level (host out) and offset of box j */
int mgx_boxset_size(const mgx_boxset* bs, int level, int64_t* npoints);
int mgx_boxset_offset(const mgx_boxset* bs, int level, int box, int64_t* offset);
/* number of overlapping box pairs (i <= j) */
int mgx_boxset_npairs(const mgx_boxset* bs, int64_t* npairs);

/* Same operators as above on packed boxes, zero outside each box. */
int mgx_box_lap(mgx_ctx* ctx, const mgx_boxset* bs, int level, int lap_type, int dtype,
    const void* in, void* out);
int mgx_box_hloc(mgx_ctx* ctx, const mgx_boxset* bs, int dtype, const void* psi,
    const double* vtot, void* hpsi);
int mgx_box_jacobi(mgx_ctx* ctx, const mgx_boxset* bs, int level, int lap_type, int dtype,
    double jacobi_factor, const void* f, void* x, void* w);
int mgx_box_restrict(mgx_ctx* ctx, const mgx_boxset* bs, int level, int dtype, const void* fine,
    void* coarse);
int mgx_box_prolong(mgx_ctx* ctx, const mgx_boxset* bs, int level, int dtype, const void* coarse,
    void* fine);
/* gamma = 0 starts the fine level from zero, as every coarse level of
 * Preconditioning::mg already does (the gamma*res start makes the V-cycle
 * indefinite on rough residuals) */
int mgx_box_precond_mg(mgx_ctx* ctx, const mgx_boxset* bs, int lap_type, int max_levels,
    double gamma, int prec_dtype, const double* res, double* out);
/* S_ij = vel*<psi_i, psi_j>; H_ij = vel*<psi_i, H psi_j> (Galerkin: hpsi on the
 * overlap plus the stencil's reach 2 points outside each box; with the cell
 * tables the value for i < j is written to both sides, without them the two
 * one-sided values are averaged: either way H is exactly symmetric); dense
 * n x n FP64, zero for pairs that do not interact.  vtot is not needed:
 * outside a box psi is zero. */
int mgx_box_gram_h(mgx_ctx* ctx, const mgx_boxset* bs, int prec, const double* psi,
    const double* hpsi, double* S, double* H);
/* res_j = sum_k psi_k theta_kj - hpsi_j on box j; res must not overlap psi or hpsi */
int mgx_box_residual(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi,
    const double* theta, double* res);
/* d_j = vel*|r_j|^2 (device n); norm = 0.5*sqrt(sum d_j Sinv_jj) host out */
int mgx_box_residual_norm(mgx_ctx* ctx, const mgx_boxset* bs, const double* res,
    const double* Sinv, double* norm);
/* Mehrstellen (FDtype = Mehrstellen, src/Control.cc:1353) on boxes:
 * Hamiltonian<T>::applyLocal's Mehrstellen branch, src/Hamiltonian.cc:106-131:
 * hpsi = -Lap_M psi + B (V psi) on each box (psi zero outside, V gathered), FP64 */
int mgx_box_hloc_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* vtot,
    double* hpsi);
/* LocGridOrbitals::computeGram / computeLocalProduct / computeMatB,
 * src/LocGridOrbitals.cc:1428-1604, with Mehrstellen operators: S_ij = vel <psi_i, psi_j>,
 * H_ij = vel <psi_i, H psi_j>, B_ij = vel <psi_i, B psi_j> (column-major, entry (i, j));
 * H psi_j and B psi_j include their one-point reach outside box j (faces, and the
 * edges of -Lap_M), so H is the Galerkin matrix; it is not symmetric (B V is not) */
int mgx_box_gram_hb_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi,
    const double* vtot, double* S, double* H, double* B);
/* MGmol::computeResidualUsingHPhi with applyB, src/MGmol.cc:1227-1291, on boxes:
 * res_j = sum_a (B psi_a) theta(a, j) - hpsi_j on box j, B psi_a with its reach */
int mgx_box_residual_mehr(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* hpsi,
    const double* theta, double* res);
/* Electron density rho(r) = sum_ij dm_ij psi_i(r) psi_j(r) on the level-0 periodic
 * grid (device, gdim[0]*gdim[1]*gdim[2] FP64, x slowest, z fastest; overwritten);
 * dm: n x n column-major, n = number of boxes.  *charge (host, may be NULL) =
 * vel * sum rho.  nel > 0 rescales rho to nel electrons and fails with
 * MGX_ENUMERIC (rho unscaled) when nel / charge is off 1 by more than 2%.
 * Needs the box set's cell tables (grid edges multiples of 8, at most 32 boxes
 * covering any 8^3 cell), else MGX_EINVAL.
 * Replaces Rho<LocGridOrbitals>::computeRhoSubdomainUsingBlas3 (src/Rho.cc:360-449)
 * with computeTotalCharge / rescaleTotalCharge (src/Rho.cc:122-208). */
int mgx_box_density(mgx_ctx* ctx, const mgx_boxset* bs, const double* psi, const double* dm,
    int n, double nel, double* rho, double* charge);

/* ------------------------------------------------ learned preconditioner ---- */
/* No counterpart in the reference: a 4-layer 3D CNN (1->16->16->16->1 channels,
 * 3^3 taps, zero padding per box, ReLU, scalar skip y = s x + conv(h)) that
 * maps a box's residual to a correction in place of the V-cycle of
 * OrbitalsPreconditioning::precond_mg (src/OrbitalsPreconditioning.cc:87-117),
 * which stays its judge and fallback.  BF16 weights/activations, FP32
 * accumulation on MFMA; FP32 master parameters trained with Adam.  Samples
 * are cubes of edge `edge`, packed one after another. */
typedef struct mgx_cnn mgx_cnn;
int mgx_cnn_nparams(void);
int mgx_cnn_create(mgx_ctx* ctx, int edge, int max_batch, mgx_cnn** out);
int mgx_cnn_destroy(mgx_cnn* c);
/* FP32 host parameters: w0[27][16][1] b0[16] w1[27][16][16] b1[16]
 * w2[27][16][16] b2[16] w3[27][1][16] b3[1] s; tap = (dx*3+dy)*3+dz */
int mgx_cnn_set_params(mgx_cnn* c, const float* host);
int mgx_cnn_get_params(mgx_cnn* c, float* host);
/* y = net(x) for nb samples (FP32, device) */
int mgx_cnn_forward(mgx_cnn* c, const float* x, int nb, float* y);
/* mean squared error of net(x) against t and its gradient (host, nparams) */
int mgx_cnn_grad(mgx_cnn* c, const float* x, const float* t, int nb, double* loss, float* grad_host);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on a batch; loss before the step */
int mgx_cnn_train_step(mgx_cnn* c, const float* x, const float* t, int nb, float lr, double* loss);
/* out_j = rms_j * net(R_j / rms_j) over nbox boxes of the CNN's edge (FP64, device) */
int mgx_cnn_apply(mgx_cnn* c, int nbox, const double* R, double* out);
/* training pairs x = R_j / rms_j, t = KR_j / rms_j (FP32, device) */
int mgx_cnn_make_pairs(mgx_cnn* c, int nbox, const double* R, const double* KR, float* x, float* t);

/* ------------------------------------------------------------- solver ---- */
/* ABPG<T>::updateWF / update_states with DFTsolver's per-step normalisation,
 * src/ABPG.cc:41-140, src/DFTsolver.cc:262-401, on a box set. */
typedef struct
{
    int lap_type;        /* MGX_LAP_4TH */
    int mg_levels;       /* preconditioner levels - 1 (Control::getMGlevels) */
    int precond_dtype;   /* MGX_F64 (plain) or MGX_F32 */
    int gram_prec;       /* MGX_PREC_F64 or MGX_PREC_SPLIT */
    double step_length;  /* Quench.step_length (1.5 for 4th order) */
    int ritz_every;      /* recompute lowest Ritz value every k steps (0 = first only) */
    int precond_start;   /* MGX_START_GAMMA (reference) or MGX_START_ZERO */
    int profile;         /* 1: time kernel groups with hipEvents (mgx_solver_profile) */
    int anderson_m;      /* ABPG.m: 0 = plain preconditioned steps (PSD), 1 = Anderson
                            extrapolation with one previous step (src/AndersonMix.cc) */
    double anderson_beta;/* ABPG.beta (1.0) */
    int conv_criterion;  /* Quench.conv_criterion: MGX_CONV_DELTAE or MGX_CONV_RESIDUAL */
    int precond_kind;    /* MGX_PRECOND_VCYCLE, or MGX_PRECOND_LEARNED (mgx_solver_set_cnn) */
    int inverse_mode;    /* S^-1 after the first step: MGX_INV_NS_F64 (Newton-Schulz refresh,
                            FP64 products), MGX_INV_NS_BF16 (its correction as one BF16 hi/lo
                            product, no counterpart in the reference) or MGX_INV_CHOL (Cholesky +
                            inverse every step, GramMatrix::updateLS / computeInverse) */
    double ns_tol;       /* Newton-Schulz target max|I - S^-1 S|; 0 = 1e-15 with FP64 S,
                            1e-10 with the split S */
    int ns_max_spec;     /* most refresh steps taken before the host checks them (0 = 4) */
} mgx_solver_cfg_t;

#define MGX_INV_NS_F64 0
#define MGX_INV_NS_BF16 1
#define MGX_INV_CHOL 2

/* MGX_PRECOND_LEARNED: a trial takes psi += alpha * CNN(R) as a plain
 * preconditioned step (Anderson history cleared) from the current state;
 * the next step evaluates it, returns to the saved state and takes the
 * V-cycle's plain step from there; the step after that evaluates the V-cycle
 * step and keeps whichever of the two reached the smaller residual norm
 * (both evaluations count as iterations).  The V-cycle alone then runs for
 * MGX_LEARNED_BACKOFF steps, a number that doubles with each rejected trial
 * in a row (at most MGX_LEARNED_BACKOFF_MAX) and returns to
 * MGX_LEARNED_BACKOFF after an accepted one. */
#define MGX_PRECOND_VCYCLE 0
#define MGX_PRECOND_LEARNED 1
#define MGX_LEARNED_BACKOFF 4
#define MGX_LEARNED_BACKOFF_MAX 256

#define MGX_START_GAMMA 0
#define MGX_START_ZERO 1

/* DFTsolver::checkConvergenceEnergy (src/DFTsolver.cc:138-167): after step 2,
 * converged when max(|dE_k|, |dE_k-1|) of the band energy is below tol (Ha);
 * or ABPG::updateWF's residual test |R| < tol (Ha). */
#define MGX_CONV_DELTAE 0
#define MGX_CONV_RESIDUAL 1

/* kernel groups timed when cfg.profile = 1.  Phase groups (several kernels):
 * GRAM, DENSE, RESID, PRECOND, NORMALIZE.  The others bracket one kernel (or
 * one kind of launch) each and also accumulate its algorithmic work: HBM bytes,
 * or flops for LEARNED, NSGEMM and NSGEMM16 (mgx_solver_profile_work). */
#define MGX_PROF_HLOC 0      /* H psi (one marching kernel) */
#define MGX_PROF_GRAM 1      /* S, H pair contractions */
#define MGX_PROF_DENSE 2     /* Cholesky, inverse, theta, trace, Ritz */
#define MGX_PROF_RESID 3     /* residual + norm */
#define MGX_PROF_JACOBI0 4   /* fine-level Jacobi sweeps without the Anderson dots */
#define MGX_PROF_PRECOND 5   /* whole V-cycle incl. update */
#define MGX_PROF_NORMALIZE 6 /* orbital normalisation */
#define MGX_PROF_LEARNED 7   /* CNN inference of the learned preconditioner */
#define MGX_PROF_CELLGRAM 8  /* the S / H cell-partials kernel */
#define MGX_PROF_SKIN 9      /* the Galerkin skin kernel (stencil reach outside each box) */
#define MGX_PROF_RESIDK 10   /* the cell residual kernel */
#define MGX_PROF_UPDATE 11   /* the orbital update (Anderson or preconditioned step) */
#define MGX_PROF_JACOBI0D 12 /* the last fine sweep forming the Anderson dots */
#define MGX_PROF_FIRST 13    /* the zero-start first fine sweep (pointwise, with the cast) */
#define MGX_PROF_NSGEMM 14   /* Newton-Schulz FP64 dense products */
#define MGX_PROF_SPMM 15     /* pair-sparse products (X S, S^-1 H) */
#define MGX_PROF_TRANSFER 16 /* fine-level restriction and prolongation */
#define MGX_PROF_NSGEMM16 17 /* Newton-Schulz BF16 correction products */
#define MGX_PROF_N 18

typedef struct
{
    int iterations;
    int converged;
    double residual;     /* Ha */
    double band_energy;  /* tr(S^-1 H), Ha */
    double seconds;
    double lowest_eig;   /* Ry */
    int learned_steps;   /* CNN trials (accepted or not), cumulative */
    int fallbacks;       /* CNN trials whose step lost to the V-cycle step from the same state */
    int ns_steps;        /* Newton-Schulz refresh steps of S^-1 (one dgemm each), cumulative */
    int chol_calls;      /* Cholesky + inverse of S (first step and refresh failures), cumulative */
    int dots_sweeps;     /* V-cycles whose last fine sweep also formed the Anderson dots, cumulative */
    int ns_continued;    /* refreshes whose unchecked steps fell short and were continued with checks */
    int ns_restarts;     /* refreshes that diverged and were replaced by Cholesky + inverse */
} mgx_solver_stats_t;

int mgx_solver_create(mgx_ctx* ctx, const mgx_boxset* bs, const mgx_solver_cfg_t* cfg,
    const double* vtot_host, const double* psi0_host, mgx_solver** s);
int mgx_solver_destroy(mgx_solver* s);
/* one wave-function step; returns residual / energy of the step's input psi */
int mgx_solver_step(mgx_solver* s, mgx_solver_stats_t* st);
int mgx_solver_run(mgx_solver* s, double tol, int maxit, mgx_solver_stats_t* st);
/* restart from the initial orbitals (kept on the device) */
int mgx_solver_reset(mgx_solver* s);
/* accumulated milliseconds and launch counts per MGX_PROF_* group (host
 * arrays of MGX_PROF_N); reset=1 clears the accumulators afterwards */
int mgx_solver_profile(mgx_solver* s, double* ms, int64_t* launches, int reset);
/* the algorithmic work accumulated per group (host array of MGX_PROF_N; bytes or
 * flops, see above), read before a reset of mgx_solver_profile */
int mgx_solver_profile_work(mgx_solver* s, double* work);
/* switch the kernel-group timing (cfg.profile) on or off between steps */
int mgx_solver_set_profile(mgx_solver* s, int on);
int mgx_solver_get_psi(mgx_solver* s, double* psi_host);
/* device pointers of the solver's psi / hpsi / S / H / residual (for inspection; the
 * residual R = psi theta - H psi of the last evaluated step is kept only with
 * mgx_solver_set_keep_residual(s, 1): by default a step whose R only feeds the
 * FP32/FP64 V-cycle from zero writes the V-cycle's start in its place) */
int mgx_solver_buffers(mgx_solver* s, double** psi, double** hpsi, double** S, double** H, double** res);
/* keep R in the residual buffer every step (on = 1).  Off (the default), the
 * residual kernel of a step whose R is read by nothing but the V-cycle (FP32 or
 * FP64 storage, zero start, no network trial around the step) writes the
 * V-cycle's start f = R, x1 = jf f in its storage type instead of R, and the
 * V-cycle skips its pointwise first sweep: one pass over the orbital points less
 * per step, the same values.  No counterpart in the reference. */
int mgx_solver_set_keep_residual(mgx_solver* s, int on);
/* device pointer of the solver's S^-1 (n x n, column-major): the inverse of the
 * S of the last evaluated step (refreshed, or Cholesky) */
int mgx_solver_inverse(mgx_solver* s, double** Sinv);
/* replace the orbitals (host, the box set's packed layout), e.g. extrapolated
 * orbitals after an ionic move (MGmol::extrapolate_orbitals keeps the solver
 * state); the Anderson history is cleared, S^-1 stays as the start of the next
 * refresh (which falls back to Cholesky when it is too far off) */
int mgx_solver_set_psi(mgx_solver* s, const double* psi_host);
/* Run each step's launch sequences as HIP graphs (on = 1): the sequence from
 * H psi to the read of the step's energy and residual norm, and the one from
 * the preconditioner to the norms of the updated psi, each captured the
 * second time the same host decisions shape it (same Newton-Schulz step
 * count, Anderson state and buffers) and replayed from then on.  Steps with
 * Cholesky, a network trial or a restore, and profiled steps, run eagerly.
 * Results are those of the eager launches.  No counterpart in the reference
 * (its steps are host loops). */
int mgx_solver_set_graph(mgx_solver* s, int on);
/* graph launches and captures so far, and graphs held (any pointer may be NULL) */
int mgx_solver_graph_stats(mgx_solver* s, int64_t* launches, int64_t* captures, int64_t* cached);
/* attach a trained network (uniform box edge equal to the network's) for
 * cfg.precond_kind = MGX_PRECOND_LEARNED; the solver does not own it */
int mgx_solver_set_cnn(mgx_solver* s, mgx_cnn* cnn);
/* Density of the solver's current (normalised) orbitals with the density matrix
 * of fully occupied non-orthogonal orbitals, dm = 2 S^-1 (S recomputed for the
 * current psi; ProjectedMatrices::setDMto2InvS, src/ProjectedMatrices.cc:265-274),
 * then as mgx_box_density.  Overwrites the solver's S, H, S^-1 and theta, which the
 * next step recomputes. */
int mgx_solver_density(mgx_solver* s, double nel, double* rho, double* charge);

#ifdef __cplusplus
}
#endif
#endif
