/*
 * oracle.h -- CPU FP64/FP32 restatement of MGmol's orbital-iteration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * (mgmol_amd/csrc).  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product path never calls it.
 *
 * Two families of functions:
 *   or_g_*  "grid" family: ghosted blocks of nfunc functions laid out exactly as
 *           MGmol's pb::GridFuncVector (x slowest, z fastest, ghost width ng,
 *           function stride sizeg).  Each follows one reference kernel, cited
 *           at the definition in oracle.cpp.
 *   or_b_*  "box" family: packed per-orbital cubic boxes on a periodic global
 *           grid, zero outside the box.  A box is the reference's masked
 *           formulation with a 0/1 cube mask (Map2Masks with a cube instead of
 *           a sphere); tests pin or_b_* against or_g_* + masks by embedding.
 *
 * Scalar suffix: _f32 = float data, _f64 = double data.  All arithmetic that
 * the reference performs in double is performed in double here too.
 * Return value: 0 on success, negative on invalid arguments.
 */
#ifndef MGX_ORACLE_H
#define MGX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    int dim[3];  /* interior points per direction (x slowest) */
    int ghosts;  /* ghost width */
    double h[3]; /* mesh spacing */
} or_grid_t;

/* box table: cubic boxes on a periodic grid of gdim points (level 0) */
typedef struct
{
    int nbox;
    int gdim[3];           /* global grid, level 0 */
    double h[3];           /* spacing, level 0 */
    const int* origin;     /* 3*nbox, level-0 indices, multiples of 2^(L-1) */
    const int* edge;       /* nbox, multiples of 2^(L-1), 2*edge+4 <= gdim */
} or_boxes_t;

/* ---------------- grid family (reference layout, ghosted) ---------------- */
int or_g_trade_f32(const or_grid_t* g, float* u, long nfunc, const int bc[3]);
int or_g_trade_f64(const or_grid_t* g, double* u, long nfunc, const int bc[3]);

int or_g_lap_f32(const or_grid_t* g, int lap_type, const float* in, float* out, long nfunc);
int or_g_lap_f64(const or_grid_t* g, int lap_type, const double* in, double* out, long nfunc);

int or_g_rhs_mehr_f32(const or_grid_t* g, const float* in, float* out, int rhs_ghosts, long nfunc);
int or_g_rhs_mehr_f64(const or_grid_t* g, const double* in, double* out, int rhs_ghosts, long nfunc);

int or_g_lap4_withpot_f32(const or_grid_t* g, const float* in, const double* pot, float* out, long ld, long nfunc);
int or_g_lap4_withpot_f64(const or_grid_t* g, const double* in, const double* pot, double* out, long ld, long nfunc);

int or_g_pointwise_f32(const or_grid_t* g, const float* a, const double* b, float* out, long nfunc);
int or_g_pointwise_f64(const or_grid_t* g, const double* a, const double* b, double* out, long nfunc);

/* Hamiltonian::applyLocal (lap_type 2: 4th order with potential; 0: Mehrstellen) */
int or_g_hloc_f64(const or_grid_t* g, int lap_type, const double* phi, const double* vtot,
    double* hphi, long ld, long nfunc, const int bc[3]);

int or_axpy_f32(long n, double alpha, const float* x, float* y);
int or_axpy_f64(long n, double alpha, const double* x, double* y);

int or_g_jacobi_f32(const or_grid_t* g, int lap_type, const float* f, float* x, float* w,
    double jacobi_factor, long nfunc, const int bc[3]);
int or_g_jacobi_f64(const or_grid_t* g, int lap_type, const double* f, double* x, double* w,
    double jacobi_factor, long nfunc, const int bc[3]);

int or_g_restrict_f32(const or_grid_t* fine, const float* uf, const or_grid_t* coarse, float* uc, long nfunc);
int or_g_restrict_f64(const or_grid_t* fine, const double* uf, const or_grid_t* coarse, double* uc, long nfunc);
int or_g_extend_f32(const or_grid_t* coarse, const float* uc, const or_grid_t* fine, float* uf, long nfunc);
int or_g_extend_f64(const or_grid_t* coarse, const double* uc, const or_grid_t* fine, double* uf, long nfunc);

/* mask: one uint8 per interior point per function (0/1), multiplies interior */
int or_g_mask_f32(const or_grid_t* g, float* u, const uint8_t* mask, long nfunc);
int or_g_mask_f64(const or_grid_t* g, double* u, const uint8_t* mask, long nfunc);

/* jacobi factors and diagonals of Laph2/Laph4/Laph4M */
double or_diag_el(int lap_type, const double h[3]);
double or_jacobi_factor(int lap_type, const double h[3]);

/* Preconditioning::mg V-cycle; masks[l] is NULL or nfunc*size_l bytes */
int or_g_mg_f32(const or_grid_t* g0, int lap_type, int max_levels, float* v, const float* f,
    long nfunc, const int bc[3], const uint8_t* const* masks);
int or_g_mg_f64(const or_grid_t* g0, int lap_type, int max_levels, double* v, const double* f,
    long nfunc, const int bc[3], const uint8_t* const* masks);

/* OrbitalsPreconditioning::precond_mg: res packed (ld) in, out packed (ld) */
int or_g_precond_f32(const or_grid_t* g0, int lap_type, int max_levels, double gamma,
    const double* res, double* out, long ld, long nfunc, const int bc[3],
    const uint8_t* const* masks);
int or_g_precond_f64(const or_grid_t* g0, int lap_type, int max_levels, double gamma,
    const double* res, double* out, long ld, long nfunc, const int bc[3],
    const uint8_t* const* masks);

double or_gamma(const double h[3], int mg_levels, double vmax, double small_eig);

/* LocalMatrices::syrk / gemm per slab (ld rows, loc_numpt per slab) */
int or_local_gram(const double* psi, long ld, long loc_numpt, int n, int subdivx, double vel, double* ss);
int or_local_product(const double* a, const double* b, long ld, long loc_numpt, int n, int subdivx,
    double vel, double* ss);
/* LocalMatrices2DistMatrix::convert followed by a dense sum */
int or_local_to_global(const double* ss, const int* gids, int subdivx, int m, int numst, double tol,
    double* global);
/* LocGridOrbitals::multiplyByMatrix + axpy(-1,hphi): res = psi*theta_loc - hphi */
int or_residual(const double* psi, const double* hphi, const double* theta, long ld, long numpt,
    int n, double* res);

/* ---------------- box family ---------------- */
long or_b_size(const or_boxes_t* b, int level);          /* total points at level */
long or_b_offset(const or_boxes_t* b, int level, int j);  /* offset of box j */

int or_b_lap_f32(const or_boxes_t* b, int level, int lap_type, const float* in, float* out);
int or_b_lap_f64(const or_boxes_t* b, int level, int lap_type, const double* in, double* out);
int or_b_hloc_f32(const or_boxes_t* b, const double* vglob, const float* in, float* out);
int or_b_hloc_f64(const or_boxes_t* b, const double* vglob, const double* in, double* out);
int or_b_jacobi_f32(const or_boxes_t* b, int level, int lap_type, const float* f, float* x, float* w, double jf);
int or_b_jacobi_f64(const or_boxes_t* b, int level, int lap_type, const double* f, double* x, double* w, double jf);
int or_b_restrict_f32(const or_boxes_t* b, int level, const float* fine, float* coarse);
int or_b_restrict_f64(const or_boxes_t* b, int level, const double* fine, double* coarse);
int or_b_prolong_f32(const or_boxes_t* b, int level, const float* coarse, float* fine);
int or_b_prolong_f64(const or_boxes_t* b, int level, const double* coarse, double* fine);
int or_b_mg_f32(const or_boxes_t* b, int lap_type, int max_levels, float* v, const float* f);
int or_b_mg_f64(const or_boxes_t* b, int lap_type, int max_levels, double* v, const double* f);
int or_b_precond_f32(const or_boxes_t* b, int lap_type, int max_levels, double gamma, const double* res, double* out);
int or_b_precond_f64(const or_boxes_t* b, int lap_type, int max_levels, double gamma, const double* res, double* out);

/* interaction pair list: pairs (i<=j) whose boxes intersect or lie within the
 * stencil reach (2 points); returns count, fills
 * up to cap entries of pi/pj (pass cap=0 to count) */
long or_b_pairs(const or_boxes_t* b, long cap, int* pi, int* pj);
/* S_ij = h^3 <psi_i, psi_j>, H_ij = h^3 <psi_i, H psi_j> (Galerkin: masked H psi on
 * the overlap plus the stencil reach outside each box), dense n x n */
int or_b_gram_h(const or_boxes_t* b, const double* psi, const double* hpsi, double* S, double* H);
/* R_j = sum_k psi_k theta_kj - hpsi_j on box j (theta column-major n x n) */
int or_b_residual(const or_boxes_t* b, const double* psi, const double* hpsi, const double* theta, double* res);
/* d_j = h^3 sum r_j^2 */
int or_b_diag_dots(const or_boxes_t* b, const double* r, double* d);
/* psi += alpha * x */
int or_b_axpy(long n, double alpha, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
