// solver.hip -- the orbital iteration on a box set, host C++ over the
// device kernels.  One step follows ABPG<T>::updateWF (src/ABPG.cc:41-68)
// inside DFTsolver::solve (src/DFTsolver.cc:262-401):
//
//   hpsi = H psi                        Hamiltonian::applyLocal
//   S, H = psi^T psi, psi^T H psi       computeGram / addHlocalij
//   L, S^-1 = chol / inverse            GramMatrix::updateLS / computeInverse
//   theta = S^-1 H, E = tr(H S^-1)      updateTheta / getEigSum
//   R = psi theta - hpsi, |R|           computeResidualUsingHPhi / ConstraintResidual
//   stop if |R| < tol                   ABPG::updateWF early return
//   psi += 1/2 step K(R)                precond_mg + update_states
//   psi_j /= |psi_j|                    LocGridOrbitals::normalize
//
// S and H are formed in one pass over psi and hpsi (S of the normalised psi
// equals the S the reference recomputes after normalize()).
#include "common.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace mgx
{
int box_hloc(mgx_ctx*, const mgx_boxset*, int, const void*, const double*, void*);
int box_gram_h(mgx_ctx*, const mgx_boxset*, int, const double*, const double*, double*, double*, double*, double*);
PairTiles boxset_tiles(const mgx_boxset*);
int64_t boxset_npairs(const mgx_boxset*);
int inverse_update_pairs(mgx_ctx*, int, const double*, const double*, const PairTiles&, double*, double*, double*,
    int*, double, bool);
int box_residual(mgx_ctx*, const mgx_boxset*, const double*, const double*, const double*, double*, double*);
int box_sumsq(mgx_ctx*, const mgx_boxset*, const double*, double*);
int box_norm(mgx_ctx*, const mgx_boxset*, const double*, const double*, double*);
int box_normalize(mgx_ctx*, const mgx_boxset*, double*, double*);
int box_hloc_normalize(mgx_ctx*, const mgx_boxset*, double*, const double*, double*, const double*);
int box_scale(mgx_ctx*, const mgx_boxset*, double*, const double*);
int box_precond(mgx_ctx*, const mgx_boxset*, int, int, double, int, const double*, double*, double*, double, bool);
int box_residual_start(mgx_ctx*, const mgx_boxset*, const double*, const double*, const double*, double*, int, int,
    int);
bool boxset_has_cells(const mgx_boxset*);
int64_t box_size(const mgx_boxset*, int);
void boxset_geom(const mgx_boxset*, int gdim[3], double h[3], int* nbox);
int chol_inv(mgx_ctx*, int, const double*, double*, double*);
int theta(mgx_ctx*, int, const double*, const double*, double*);
int trace_prod(mgx_ctx*, int, const double*, const double*, double*);
int diag_sum(mgx_ctx*, int, const double*, double*);
int inverse_update(mgx_ctx*, int, const double*, double*, double*, double*, int*);
int ritz(mgx_ctx*, int, const double*, const double*, double*, double*);
int box_anderson_dots(mgx_ctx*, const mgx_boxset*, int, const void*, double, const void*, const double*, double*,
    double*);
int box_anderson_update_sq(mgx_ctx*, const mgx_boxset*, int, const double*, double*, const void*, double,
    const void*, void*, const double*, double, double*);
int box_psi_step_sq(mgx_ctx*, const mgx_boxset*, int, double*, const void*, double, double*);
int box_precond_x(mgx_ctx*, const mgx_boxset*, int, int, double, int, const double*, void*, const void*, double,
    double*, bool*, const double*, bool);
int box_anderson_dots_total(mgx_ctx*, const mgx_boxset*, const double*, const double*, double*);
int inverse_update_spec(mgx_ctx*, int, const double*, const PairTiles&, double*, double*, int, bool, double*, double*,
    bool);
int ns_spec_check(const double*, int, double, bool, int*);
int cnn_apply_boxes(mgx_cnn*, int, int, const double*, void*, int);
bool boxset_uniform_edge(const mgx_boxset*, int*);
int boxset_maxedge(const mgx_boxset*);
int box_anderson_update_nrm(mgx_ctx*, const mgx_boxset*, int, const double*, double*, const void*, double,
    const void*, void*, const double*, double, double*);
int box_density(mgx_ctx*, const mgx_boxset*, const double*, const double*, int, double, double*, double*);
int box_hloc_mehr(mgx_ctx*, const mgx_boxset*, const double*, const double*, double*);
int box_bpsi_mehr(mgx_ctx*, const mgx_boxset*, const double*, double*);
int box_gram_hb_mehr(mgx_ctx*, const mgx_boxset*, const double*, const double*, const double*, const double*, double*,
    double*, double*);
int box_residual_mehr(mgx_ctx*, const mgx_boxset*, const double*, const double*, const double*, const double*,
    double*);

// y = a x (n values)
__global__ void k_scaled_copy(const double* x, int64_t n, double a, double* y)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        y[t] = a * x[t];
}
// what a tail sequence leaves for the host to record (the same on every
// replay of its graph)
struct TailOut
{
    bool norms_done = false;  // d holds the norms of the updated psi
    bool normalized = false;  // psi is normalised already
    bool anderson   = false;  // the Anderson update ran: psi <-> x0 swap
    bool keep       = false;  // and the V-cycle's correction becomes the history (kr0 <-> kr1)
    bool dots_total = false;  // the V-cycle's last sweep formed the Anderson dots
};

}  // namespace mgx

struct mgx_solver
{
    mgx_ctx* ctx          = nullptr;
    const mgx_boxset* bs  = nullptr;
    mgx_solver_cfg_t cfg{};
    int n          = 0;
    int64_t npts   = 0;
    double h[3]    = { 0, 0, 0 };
    double vmax    = 0.;
    double lowest  = 0.;
    int iter       = 0;
    double *psi = nullptr, *psi0 = nullptr, *hpsi = nullptr, *res = nullptr, *vtot = nullptr;
    MgxProf prof;
    double *S = nullptr, *H = nullptr, *L = nullptr, *Sinv = nullptr, *th = nullptr;
    double *eig = nullptr, *d = nullptr, *scal = nullptr, *nswork = nullptr;
    double *Sv = nullptr, *Hv = nullptr;  // S, H per interacting pair
    // Mehrstellen: B psi, the B-matrix and its inverse
    double *bpsi = nullptr, *Bm = nullptr, *invB = nullptr;
    double* x0 = nullptr;  // Anderson history: previous psi
    void* kr0  = nullptr;  // and previous correction, in the preconditioner's type
    void* kr1  = nullptr;  // the V-cycle's correction (swapped with kr0 after the update)
    void* kcnn = nullptr;  // the network's correction (learned steps)
    // DFTsolver::checkConvergenceEnergy state (band energy, Ha)
    double e_prev = 0., deig = 1e9, deig2 = 1e9;
    bool accelerate = false;
    int mm          = 0;  // Anderson history length in use
    // the normalisation of the last update is pending: d holds vel*|psi_j|^2
    // of the unnormalised psi; the next Hpsi applies it (or flush_norm)
    bool pending_norm = false;
    // the residual kernel writes the V-cycle's zero start instead of R this
    // step (set per step); keep_res: always write R (a caller reads it)
    bool start_fused = false;
    bool keep_res    = false;
    // learned preconditioner: the state before the last CNN step, to undo it
    mgx_cnn* cnn       = nullptr;
    double *psi_prev = nullptr, *r_prev = nullptr;  // the saved state of a trial
    double *psi_cnn = nullptr, *r_cnn = nullptr;    // the network's step from it, evaluated
    int trial         = 0;  // 0: none, 1: network step taken, 2: V-cycle step from the same state taken
    double r_cnn_norm = 0.;
    int backoff        = 0;
    int backoff_len    = MGX_LEARNED_BACKOFF;
    int learned_steps = 0, fallbacks = 0;
    mgx::PairTiles tiles;
    bool have_inv = false;  // Sinv holds the inverse of a recent S
    // the inverse of the step before the last refresh (FP64 refresh): the
    // next refresh starts from 2 Sinv - xprev when have_xprev
    double* xprev   = nullptr;
    bool have_xprev = false;
    int64_t ns_steps = 0, chol_calls = 0, dots_sweeps = 0;
    int ns_pred = 2;  // Newton-Schulz steps taken without host checks (the last step's need)
    int ns_max_spec = 4;
    double ns_tol   = 1e-15;
    int last_more   = 0;  // outcome of the last ns_spec_check (0 accepted, 1 continued, -1 Cholesky)
    int64_t ns_continued = 0, ns_restarts = 0;  // checked continuations, Cholesky restarts of the refresh
    int64_t gsize = 0;
    double* host = nullptr;  // pinned: [energy, norm]
    // HIP graphs of the step's launch sequences, by the host decisions that
    // shape them (mgx_solver_set_graph)
    struct Graph
    {
        hipGraphExec_t exec = nullptr;
        uint64_t ws_gen     = 0;  // the context's workspace generation at capture
        int seen            = 0;
        bool failed         = false;
        mgx::TailOut out;
    };
    bool use_graph = false;
    std::unordered_map<uint64_t, Graph> graphs;
    int64_t graph_launches = 0, graph_captures = 0;
};

namespace mgx
{
namespace
{

void free_solver(mgx_solver* s)
{
    double* bufs[] = { s->psi, s->psi0, s->hpsi, s->res, s->vtot, s->S, s->H, s->L, s->Sinv, s->th, s->eig, s->d,
        s->scal, s->nswork, s->Sv, s->Hv, s->x0, s->psi_prev, s->r_prev, s->psi_cnn, s->r_cnn, s->bpsi, s->Bm,
        s->invB, s->xprev };
    for (double* b : bufs)
        if (b) hipFree(b);
    if (s->kr0) (void)hipFree(s->kr0);
    if (s->kr1) (void)hipFree(s->kr1);
    if (s->kcnn) (void)hipFree(s->kcnn);
    if (s->host) hipHostFree(s->host);
    for (auto& kv : s->graphs)
        if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
    s->graphs.clear();
    for (hipEvent_t e : s->prof.pool)
        hipEventDestroy(e);
    s->prof.pool.clear();
}

// theta, the band energy, the residual and its norm from the current S^-1,
// and their copy to the host (read after the caller's stream wait); tk is
// the open DENSE timing token
int project(mgx_solver* s, int spec, bool bf16, int& tk)
{
    mgx_ctx* ctx = s->ctx;
    const int n  = s->n;
    int r;
    // theta = S^-1 H with H pair-sparse
    if ((r = spmm_pairs(ctx, n, s->Sinv, s->Hv, s->tiles, s->th))) return r;
    // tr(H S^-1) = tr(S^-1 H) = tr(theta)   (ProjectedMatrices::getEigSum)
    if ((r = diag_sum(ctx, n, s->th, s->scal))) return r;
    prof_end(ctx, tk);
    tk = prof_begin(ctx, MGX_PROF_RESID);
    // the residual itself, or (start_fused) the V-cycle's zero start formed
    // from it in the same kernel; the norms in d either way
    if ((r = s->start_fused ? box_residual_start(ctx, s->bs, s->psi, s->hpsi, s->th, s->d, s->cfg.precond_dtype,
                                  s->cfg.lap_type, s->cfg.mg_levels)
                            : box_residual(ctx, s->bs, s->psi, s->hpsi, s->th, s->res, s->d)))
        return r;
    if ((r = box_norm(ctx, s->bs, s->d, s->Sinv, s->scal + 1))) return r;
    prof_end(ctx, tk);
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->host, s->scal, (8 + spec + (bf16 ? 1 : 0)) * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    return MGX_OK;
}

// The launches of a step up to the host read of its scalars: H psi (with the
// pending normalisation), S and H, S^-1 (Cholesky + inverse, or spec
// Newton-Schulz steps whose deltas the host checks), theta, the energy, the
// residual and its norm.  Changes no host state.
int head_body(mgx_solver* s, bool pending, bool chol, int spec, bool bf16)
{
    mgx_ctx* ctx     = s->ctx;
    const double np  = (double)s->npts;
    int rc;
    // H psi: psi read, H psi written (FP64), V once; psi written back when
    // the pending normalisation is applied on the way
    int tk = prof_begin(ctx, MGX_PROF_HLOC, (pending ? 24. : 16.) * np + 8. * (double)s->gsize);
    if (pending)
    {
        // LocGridOrbitals::normalize of the last update, fused with Hpsi
        if ((rc = box_hloc_normalize(ctx, s->bs, s->psi, s->vtot, s->hpsi, s->d))) return rc;
    }
    else if ((rc = box_hloc(ctx, s->bs, MGX_F64, s->psi, s->vtot, s->hpsi)))
        return rc;
    prof_end(ctx, tk);
    tk = prof_begin(ctx, MGX_PROF_GRAM);
    if ((rc = box_gram_h(ctx, s->bs, s->cfg.gram_prec, s->psi, s->hpsi, s->S, s->H, s->Sv, s->Hv))) return rc;
    prof_end(ctx, tk);
    tk = prof_begin(ctx, MGX_PROF_DENSE);
    // S^-1: Cholesky + inverse (GramMatrix::updateLS / computeInverse), or
    // the Newton-Schulz refresh from the previous inverse
    if (chol)
    {
        if ((rc = chol_inv(ctx, s->n, s->S, s->L, s->Sinv))) return rc;
    }
    else if ((rc = inverse_update_spec(ctx, s->n, s->Sv, s->tiles, s->Sinv, s->nswork, spec, bf16, s->scal + 8,
                  bf16 ? nullptr : s->xprev, s->have_xprev)))
        return rc;
    return project(s, spec, bf16, tk);
}


// The launches of a step after the host's decisions: the preconditioner (or
// the network), the orbital update and the norms.  Changes no host state.
int tail_body(mgx_solver* s, bool use_cnn, bool restored, bool plain_step, double gamma, TailOut& o)
{
    mgx_ctx* ctx       = s->ctx;
    const double np    = (double)s->npts;
    const double alpha = 0.5 * s->cfg.step_length;
    int rc;
    o = TailOut();
    bool dots_done = false;  // the V-cycle formed the Anderson dots in d
    int tk         = prof_begin(ctx, MGX_PROF_PRECOND);
    const int prec = s->cfg.precond_dtype;
    // the correction's type: FP32 out of the FP16 V-cycle
    const int kprec = prec == MGX_F16 ? MGX_F32 : prec;
    if (!use_cnn && !s->accelerate && !restored && prec != MGX_F16)
    {
        // preconditioned step psi += alpha K(R) (ABPG::update_states, no
        // extrapolation), fused with the V-cycle's last kernel
        if ((rc = box_precond(ctx, s->bs, s->cfg.lap_type, s->cfg.mg_levels, gamma, prec, s->res, nullptr, s->psi,
                 alpha, s->start_fused)))
            return rc;
    }
    else
    {
        // the correction K(R) (or the network's), in the preconditioner's type
        const void* kr = nullptr;
        if (use_cnn)
        {
            int edge = 0;
            boxset_uniform_edge(s->bs, &edge);
            // 4 conv3d layers 1->16->16->16->1, 3^3 taps: 2 * 27 * 544 flops per voxel
            const int tl = prof_begin(ctx, MGX_PROF_LEARNED, 2. * 27. * 544. * np);
            if ((rc = cnn_apply_boxes(s->cnn, s->n, edge, s->res, s->kcnn, kprec))) return rc;
            prof_end(ctx, tl);
            kr = s->kcnn;
        }
        else
        {
            // d holds the residual's per-box sums unless R was just restored
            if ((rc = box_precond_x(ctx, s->bs, s->cfg.lap_type, s->cfg.mg_levels, gamma, prec, s->res, s->kr1,
                     (s->accelerate && s->mm > 0) ? s->kr0 : nullptr, alpha, s->d, &dots_done,
                     restored ? nullptr : s->d, s->start_fused && !restored)))
                return rc;
            kr = s->kr1;
        }
        const double sk = kprec == MGX_F64 ? 8. : 4.;
        if (!s->accelerate || plain_step)
        {
            // psi += alpha K(R), with the norms of the new psi: psi read and
            // written, the correction read
            const int tu = prof_begin(ctx, MGX_PROF_UPDATE, (16. + sk) * np);
            if ((rc = box_psi_step_sq(ctx, s->bs, kprec, s->psi, kr, alpha, s->d))) return rc;
            prof_end(ctx, tu);
            o.norms_done = true;
        }
        else
        {
            // AndersonMix::update with m = 1 on f = alpha K(R); theta is
            // formed on the device with the dot products (no host wait)
            const double* thp = nullptr;
            double beta       = 1.;
            if (s->mm > 0)
            {
                if (dots_done)
                {
                    rc           = box_anderson_dots_total(ctx, s->bs, s->Sinv, s->d, s->scal + 2);
                    o.dots_total = true;
                }
                else
                    rc = box_anderson_dots(ctx, s->bs, kprec, kr, alpha, s->kr0, s->Sinv, s->d, s->scal + 2);
                if (rc) return rc;
                thp  = s->scal + 4;
                beta = s->cfg.anderson_beta;
            }
            // the new psi goes to the history buffer, and the buffers swap:
            // the old psi is the next step's x0; the norms of the new psi too.
            // The V-cycle's correction buffer becomes the history (kr0, kr1
            // swap); the network's correction is copied into kr0 instead
            // boxes up to 32^3: the normalisation in the same pass
            o.keep       = kr == s->kr1;
            o.normalized = boxset_maxedge(s->bs) <= 32;
            // psi and the previous psi read, the new psi written (FP64); the
            // correction and the previous one read (+ the copy when not kept)
            const int tu = prof_begin(ctx, MGX_PROF_UPDATE, (24. + (o.keep ? 2. : 3.) * sk) * np);
            if ((rc = o.normalized ? box_anderson_update_nrm(ctx, s->bs, kprec, s->psi, s->x0, kr, alpha, s->kr0,
                                         o.keep ? nullptr : s->kr0, thp, beta, s->d)
                                   : box_anderson_update_sq(ctx, s->bs, kprec, s->psi, s->x0, kr, alpha, s->kr0,
                                         o.keep ? nullptr : s->kr0, thp, beta, s->d)))
                return rc;
            prof_end(ctx, tu);
            o.anderson   = true;
            o.norms_done = true;
        }
    }
    prof_end(ctx, tk);
    // LocGridOrbitals::normalize: the norms now (fused into the update when
    // it was not the V-cycle's fused step), the scaling with the next Hpsi
    // unless the Anderson update has applied it already
    tk = prof_begin(ctx, MGX_PROF_NORMALIZE);
    if (!o.norms_done && (rc = box_sumsq(ctx, s->bs, s->psi, s->d))) return rc;
    prof_end(ctx, tk);
    return MGX_OK;
}

// The same for the Mehrstellen discretisation (FDtype = Mehrstellen): H psi
// = -Lap_M psi + B V psi and B psi on the boxes, the Galerkin S, H and
// B-matrix, S^-1 and B-matrix^-1 by Cholesky + inverse (GramMatrix,
// ProjectedMatricesMehrstellen::computeInvB,
// src/ProjectedMatricesMehrstellen.cc:38-64), theta = B-matrix^-1 H
// (updateTheta, .h:57-62), the band energy tr(HB S^-1) = tr(theta)
// (getEigSum, src/ProjectedMatrices.cc:550-569: HB = (S theta + theta^T S) / 2),
// the residual (B Psi) theta - H Psi and its norm.  Launched eagerly.
int head_mehr(mgx_solver* s, bool pending)
{
    mgx_ctx* ctx = s->ctx;
    int rc;
    int tk = prof_begin(ctx, MGX_PROF_HLOC);
    if (pending && (rc = box_scale(ctx, s->bs, s->psi, s->d))) return rc;
    if ((rc = box_hloc_mehr(ctx, s->bs, s->psi, s->vtot, s->hpsi))) return rc;
    if ((rc = box_bpsi_mehr(ctx, s->bs, s->psi, s->bpsi))) return rc;
    prof_end(ctx, tk);
    tk = prof_begin(ctx, MGX_PROF_GRAM);
    if ((rc = box_gram_hb_mehr(ctx, s->bs, s->psi, s->hpsi, s->bpsi, s->vtot, s->S, s->H, s->Bm))) return rc;
    prof_end(ctx, tk);
    tk = prof_begin(ctx, MGX_PROF_DENSE);
    if ((rc = chol_inv(ctx, s->n, s->S, s->L, s->Sinv))) return rc;
    if ((rc = chol_inv(ctx, s->n, s->Bm, s->L, s->invB))) return rc;
    if ((rc = theta(ctx, s->n, s->invB, s->H, s->th))) return rc;
    if ((rc = diag_sum(ctx, s->n, s->th, s->scal))) return rc;
    prof_end(ctx, tk);
    tk = prof_begin(ctx, MGX_PROF_RESID);
    if ((rc = box_residual_mehr(ctx, s->bs, s->psi, s->bpsi, s->hpsi, s->th, s->res))) return rc;
    if ((rc = box_sumsq(ctx, s->bs, s->res, s->d))) return rc;
    if ((rc = box_norm(ctx, s->bs, s->d, s->Sinv, s->scal + 1))) return rc;
    prof_end(ctx, tk);
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->host, s->scal, 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return MGX_OK;
}

uint64_t mix_key(uint64_t h, uint64_t v)
{
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

// Runs body, or replays its HIP graph.  A sequence seen for the first time
// runs eagerly (its workspaces grow then); the second time it is captured
// (relaxed mode) into a graph and launched; later steps with the same key
// replay the graph.  A sequence that cannot be captured keeps running
// eagerly.  body must enqueue work on the context's stream only and change
// no host state; the TailOut it fills is stored with the graph.
template <class F>
int run_seq(mgx_solver* s, uint64_t key, TailOut* out, F&& body)
{
    mgx_ctx* ctx = s->ctx;
    if (!s->use_graph || ctx->prof) return body();
    mgx_solver::Graph& g = s->graphs[key];
    if (g.exec && g.ws_gen != ctx->ws_gen)
    {
        // a workspace the graph writes has moved since: capture again
        hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.seen = 0;
    }
    if (g.exec)
    {
        MGX_CHECK_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
        s->graph_launches++;
        if (out) *out = g.out;
        return MGX_OK;
    }
    if (g.failed || g.seen++ == 0) return body();
    MGX_CHECK_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    const int rc     = body();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc == MGX_OK && e == hipSuccess && graph
        && hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess)
    {
        hipGraphDestroy(graph);
        g.ws_gen = ctx->ws_gen;
        if (out) g.out = *out;
        s->graph_captures++;
        MGX_CHECK_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
        s->graph_launches++;
        return MGX_OK;
    }
    // not capturable: run it eagerly, now and from here on
    if (graph) hipGraphDestroy(graph);
    g.exec   = nullptr;
    g.failed = true;
    (void)hipGetLastError();
    return body();
}

int step_body(mgx_solver* s, double tol, mgx_solver_stats_t* st, bool* done)
{
    mgx_ctx* ctx = s->ctx;
    int rc, tk = -1;
    const int n = s->n;
    const double ns_tol = s->ns_tol;
    const bool bf16     = s->cfg.inverse_mode == MGX_INV_NS_BF16;
    // Cholesky on the first step (every step with MGX_INV_CHOL); else as
    // many Newton-Schulz steps as the last refresh needed, unchecked, their
    // deltas checked with the energy and the residual norm below
    const bool mehr    = s->cfg.lap_type == MGX_LAP_MEHRSTELLEN;
    const bool chol    = mehr || !s->have_inv || s->cfg.inverse_mode == MGX_INV_CHOL;
    int spec           = chol ? 0 : std::max(1, std::min(s->ns_pred, s->ns_max_spec));
    const bool pending = s->pending_norm;
    // the residual goes straight into the V-cycle's zero start unless R itself
    // is read this step (a network trial, its evaluation, or the caller)
    {
        const int prec     = s->cfg.precond_dtype;
        const bool learned = s->cfg.precond_kind == MGX_PRECOND_LEARNED;
        const bool need_r  = s->keep_res || (learned && (s->trial != 0 || s->backoff == 0));
        s->start_fused = !mehr && !need_r && boxset_has_cells(s->bs) && (prec == MGX_F32 || prec == MGX_F64)
                         && s->cfg.precond_start == MGX_START_ZERO;
    }
    if (mehr)
        rc = head_mehr(s, pending);  // Cholesky of S and of the B-matrix every step
    else if (chol)
        rc = head_body(s, pending, true, 0, bf16);  // waits for potrf's info: never captured
    else
    {
        uint64_t key = 1;
        for (uint64_t v : { (uint64_t)pending, (uint64_t)spec, (uint64_t)bf16, (uint64_t)(uintptr_t)s->psi,
                 (uint64_t)s->start_fused, (uint64_t)s->have_xprev })
            key = mix_key(key, v);
        rc = run_seq(s, key, nullptr, [&]() { return head_body(s, pending, false, spec, bf16); });
    }
    if (rc) return rc;
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->pending_norm = false;
    if (chol)
    {
        s->chol_calls++;
        s->have_inv   = true;
        s->have_xprev = false;
    }
    else if (!bf16)
        s->have_xprev = true;  // xprev holds this step's starting inverse
    if (spec > 0)
    {
        int needed     = 0;
        const int more = ns_spec_check(s->host + 8, spec, ns_tol, bf16, &needed);
        s->last_more   = more;
        if (more == 0)
        {
            s->ns_steps += spec;
            s->ns_pred = std::max(needed, 1);
        }
        else
        {
            // the unchecked steps were too few or did not converge: finish
            // the inverse as inverse_update_pairs does (Cholesky when the
            // steps diverged), then theta, the residual and its norm again
            tk = prof_begin(ctx, MGX_PROF_DENSE);
            int steps = 0;
            if (more < 0)
            {
                if ((rc = chol_inv(ctx, n, s->S, s->L, s->Sinv))) return rc;
                steps = -1;
            }
            else if ((rc = inverse_update_pairs(ctx, n, s->S, s->Sv, s->tiles, s->L, s->Sinv, s->nswork, &steps,
                          ns_tol, bf16)))
                return rc;
            s->ns_steps += spec + std::max(steps, 0);
            if (steps < 0) s->chol_calls++;
            if (more < 0 || steps < 0)
                s->ns_restarts++;
            else
                s->ns_continued++;
            s->ns_pred = steps < 0 ? 2 : spec + steps;
            spec       = 0;
            if ((rc = project(s, 0, bf16, tk))) return rc;
            MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
    }
    const double e_ry = s->host[0], rnorm = s->host[1];
    if (!std::isfinite(e_ry) || !std::isfinite(rnorm))
        return set_err(ctx, MGX_ENUMERIC, "non-finite energy or residual");
    st->band_energy = 0.5 * e_ry;  // Ry -> Ha
    st->residual    = rnorm;
    // energy history (DFTsolver::checkConvergenceEnergy): deig2 = max of the
    // last two band-energy changes, 1e9 until two steps exist
    if (s->iter > 0)
    {
        const double deig_old = s->deig;
        s->deig               = std::fabs(st->band_energy - s->e_prev);
        s->deig2              = std::max(s->deig, deig_old);
    }
    s->e_prev = st->band_energy;
    st->ns_steps     = (int)s->ns_steps;
    st->chol_calls   = (int)s->chol_calls;
    st->ns_continued = (int)s->ns_continued;
    st->ns_restarts  = (int)s->ns_restarts;
    // Anderson starts once the eigenvalue-sum change (2 x Ry = 4 x Ha of the
    // band energy) is below 1e-2 per state
    s->accelerate = s->accelerate || (s->cfg.anderson_m > 0 && 4. * s->deig2 < 1e-2 * n);
    if (s->cfg.conv_criterion == MGX_CONV_DELTAE)
        *done = (tol > 0. && s->iter > 1 && s->deig2 < tol);
    else
        *done = (tol > 0. && rnorm < tol);
    if (*done) return MGX_OK;
    // learned preconditioner.  A trial takes the network's correction as a
    // plain preconditioned step (Anderson history cleared) from the saved
    // state (psi, R); the next step evaluates it, keeps it aside and takes
    // the V-cycle's plain step from the same saved state; the step after that
    // evaluates the V-cycle step and keeps whichever of the two reached the
    // smaller residual norm.  The V-cycle alone then runs for backoff_len
    // steps; each rejection in a row doubles that (up to
    // MGX_LEARNED_BACKOFF_MAX), an accepted trial resets it.
    bool use_cnn = false, restored = false, plain_step = false;
    if (s->cfg.precond_kind == MGX_PRECOND_LEARNED)
    {
        const size_t bytes = s->npts * sizeof(double);
        if (s->trial == 1)
        {
            s->r_cnn_norm = rnorm;
            MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->psi_cnn, s->psi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->r_cnn, s->res, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->psi, s->psi_prev, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->res, s->r_prev, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            restored   = true;
            plain_step = true;
            s->trial   = 2;
        }
        else if (s->trial == 2)
        {
            s->trial = 0;
            if (s->r_cnn_norm < rnorm)
            {
                // the network's step wins: back to it
                MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->psi, s->psi_cnn, bytes, hipMemcpyDeviceToDevice, ctx->stream));
                MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->res, s->r_cnn, bytes, hipMemcpyDeviceToDevice, ctx->stream));
                restored       = true;
                s->backoff_len = MGX_LEARNED_BACKOFF;
                s->backoff     = MGX_LEARNED_BACKOFF;
            }
            else
            {
                s->fallbacks++;
                s->backoff     = s->backoff_len;
                s->backoff_len = std::min(2 * s->backoff_len, MGX_LEARNED_BACKOFF_MAX);
            }
        }
        else if (s->backoff > 0)
            s->backoff--;
        else
        {
            if (!s->cnn) return set_err(ctx, MGX_EINVAL, "no network attached (mgx_solver_set_cnn)");
            MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->psi_prev, s->psi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            MGX_CHECK_HIP(ctx, hipMemcpyAsync(s->r_prev, s->res, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            use_cnn    = true;
            plain_step = true;
            s->trial   = 1;
            s->learned_steps++;
        }
        if (plain_step || restored) s->mm = 0;
    }
    double gamma = 0.;
    if (s->cfg.precond_start == MGX_START_GAMMA)
    {
        // lowest eigenvalue of (H, S) for gamma (ProjectedMatrices::getLowestEigenvalue)
        if (s->iter == 0 || (s->cfg.ritz_every > 0 && s->iter % s->cfg.ritz_every == 0))
        {
            tk = prof_begin(ctx, MGX_PROF_DENSE);
            if ((rc = ritz(ctx, n, s->H, s->S, s->eig, nullptr))) return rc;
            prof_end(ctx, tk);
            MGX_CHECK_HIP(ctx,
                hipMemcpyAsync(s->host + 2, s->eig, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
            s->lowest = s->host[2];
        }
        gamma = mgx_gamma(s->h, s->cfg.mg_levels, s->vmax, s->lowest);
    }
    st->lowest_eig    = s->lowest;
    st->learned_steps = s->learned_steps;
    st->fallbacks     = s->fallbacks;
    TailOut o;
    if (!use_cnn && !restored)
    {
        uint64_t key = 2;
        uint64_t gbits;
        std::memcpy(&gbits, &gamma, sizeof(gbits));
        for (uint64_t v : { (uint64_t)s->accelerate, (uint64_t)plain_step, (uint64_t)(s->mm > 0), gbits,
                 (uint64_t)(uintptr_t)s->psi, (uint64_t)(uintptr_t)s->x0, (uint64_t)(uintptr_t)s->kr0,
                 (uint64_t)(uintptr_t)s->kr1, (uint64_t)s->start_fused })
            key = mix_key(key, v);
        rc = run_seq(s, key, &o, [&]() { return tail_body(s, false, false, plain_step, gamma, o); });
    }
    else
        rc = tail_body(s, use_cnn, restored, plain_step, gamma, o);
    if (rc) return rc;
    if (o.dots_total) s->dots_sweeps++;
    if (o.anderson)
    {
        std::swap(s->psi, s->x0);
        if (o.keep) std::swap(s->kr0, s->kr1);
        s->mm = std::min(s->mm + 1, s->cfg.anderson_m);
    }
    st->dots_sweeps = (int)s->dots_sweeps;
    s->pending_norm = !o.normalized;
    s->iter++;
    return MGX_OK;
}

// apply a pending normalisation, so that psi is normalised when seen from outside
int flush_norm(mgx_solver* s)
{
    if (!s->pending_norm) return MGX_OK;
    int rc = box_scale(s->ctx, s->bs, s->psi, s->d);
    if (rc) return rc;
    s->pending_norm = false;
    MGX_CHECK_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return MGX_OK;
}

int step_impl(mgx_solver* s, double tol, mgx_solver_stats_t* st, bool* done)
{
    mgx_ctx* ctx = s->ctx;
    ctx->prof    = s->cfg.profile ? &s->prof : nullptr;
    int rc       = step_body(s, tol, st, done);
    prof_collect(ctx);
    ctx->prof = nullptr;
    return rc;
}

}  // namespace
}  // namespace mgx

using namespace mgx;

extern "C" {

int mgx_solver_create(mgx_ctx* ctx, const mgx_boxset* bs, const mgx_solver_cfg_t* cfg, const double* vtot_host,
    const double* psi0_host, mgx_solver** out)
{
    MGX_REQUIRE(ctx, bs && cfg && vtot_host && psi0_host && out, "bad solver arguments");
    MGX_REQUIRE(ctx, cfg->lap_type == MGX_LAP_4TH || cfg->lap_type == MGX_LAP_MEHRSTELLEN,
        "solver supports the 4th-order and the Mehrstellen operators");
    if (cfg->lap_type == MGX_LAP_MEHRSTELLEN)
    {
        MGX_REQUIRE(ctx, cfg->precond_dtype == MGX_F32 || cfg->precond_dtype == MGX_F64,
            "the Mehrstellen solver takes an FP32 or FP64 V-cycle");
        MGX_REQUIRE(ctx, cfg->gram_prec == MGX_PREC_F64, "the Mehrstellen solver forms S, H and B in FP64");
        MGX_REQUIRE(ctx, cfg->precond_kind == MGX_PRECOND_VCYCLE, "the Mehrstellen solver takes the V-cycle");
    }
    MGX_REQUIRE(ctx, cfg->precond_dtype == MGX_F32 || cfg->precond_dtype == MGX_F64 || cfg->precond_dtype == MGX_F16,
        "bad precond dtype");
    MGX_REQUIRE(ctx, cfg->gram_prec == MGX_PREC_F64 || cfg->gram_prec == MGX_PREC_SPLIT, "bad gram precision");
    MGX_REQUIRE(ctx, cfg->precond_start == MGX_START_GAMMA || cfg->precond_start == MGX_START_ZERO,
        "bad precond start");
    MGX_REQUIRE(ctx, cfg->step_length > 0., "step length must be positive");
    MGX_REQUIRE(ctx, cfg->anderson_m == 0 || cfg->anderson_m == 1, "Anderson history length must be 0 or 1");
    MGX_REQUIRE(ctx, cfg->precond_kind == MGX_PRECOND_VCYCLE || cfg->precond_kind == MGX_PRECOND_LEARNED,
        "bad preconditioner kind");
    if (cfg->precond_kind == MGX_PRECOND_LEARNED)
    {
        int e = 0;
        MGX_REQUIRE(ctx, boxset_uniform_edge(bs, &e), "the learned preconditioner needs boxes of one edge");
    }
    MGX_REQUIRE(ctx, cfg->conv_criterion == MGX_CONV_DELTAE || cfg->conv_criterion == MGX_CONV_RESIDUAL,
        "bad convergence criterion");
    MGX_REQUIRE(ctx, cfg->inverse_mode == MGX_INV_NS_F64 || cfg->inverse_mode == MGX_INV_NS_BF16
                         || cfg->inverse_mode == MGX_INV_CHOL,
        "bad inverse mode");
    MGX_REQUIRE(ctx, cfg->ns_tol >= 0. && cfg->ns_tol < 0.25, "Newton-Schulz tolerance must be in [0, 0.25)");
    MGX_REQUIRE(ctx, cfg->ns_max_spec >= 0 && cfg->ns_max_spec <= 8, "ns_max_spec must be in [0, 8]");
    int64_t dummy = 0;
    MGX_REQUIRE(ctx, mgx_boxset_size(bs, cfg->mg_levels, &dummy) == MGX_OK, "box set has too few levels");
    int tuned = 0;
    mgx_boxset_get_tune(bs, nullptr, &tuned);
    if (!tuned)
    {
        // launch shapes of the fine-level kernels for this box set
        const int rc = mgx_boxset_autotune(ctx, const_cast<mgx_boxset*>(bs), nullptr, nullptr);
        if (rc) return rc;
    }
    mgx_solver* s = new mgx_solver();
    s->ctx        = ctx;
    s->bs         = bs;
    s->cfg        = *cfg;
    mgx_boxset_size(bs, 0, &s->npts);
    int g[3];
    boxset_geom(bs, g, s->h, &s->n);
    s->tiles = boxset_tiles(bs);
    const int64_t gsize = (int64_t)g[0] * g[1] * g[2];
    s->gsize            = gsize;
    // accuracy of the refreshed inverse: FP64 (1e-15) with FP64 S; with the
    // split S (entries good to ~1e-8) 1e-10 suffices, one step fewer
    s->ns_tol      = cfg->ns_tol > 0. ? cfg->ns_tol : (cfg->gram_prec == MGX_PREC_SPLIT ? 1e-10 : 1e-15);
    s->ns_max_spec = cfg->ns_max_spec > 0 ? cfg->ns_max_spec : 4;
    const size_t nn     = (size_t)s->n * s->n;
    auto alloc = [&](double** p, size_t cnt) { return hipMalloc((void**)p, cnt * sizeof(double)) == hipSuccess; };
    const size_t tsize = cfg->precond_dtype == MGX_F64 ? sizeof(double) : sizeof(float);
    bool ok = alloc(&s->psi, s->npts) && alloc(&s->psi0, s->npts) && alloc(&s->hpsi, s->npts) && alloc(&s->res, s->npts) && alloc(&s->vtot, gsize)
              && alloc(&s->S, nn) && alloc(&s->H, nn) && alloc(&s->L, nn) && alloc(&s->Sinv, nn) && alloc(&s->th, nn)
              && alloc(&s->eig, s->n) && alloc(&s->d, 2 * (size_t)s->n + 8) && alloc(&s->scal, 32) && alloc(&s->nswork, 2 * nn)
              && alloc(&s->Sv, (size_t)std::max<int64_t>(boxset_npairs(bs), 1))
              && alloc(&s->Hv, (size_t)std::max<int64_t>(boxset_npairs(bs), 1))
              && (cfg->anderson_m == 0
                  || (alloc(&s->x0, s->npts) && hipMalloc(&s->kr0, s->npts * tsize) == hipSuccess))
              // the V-cycle's correction is stored for Anderson, for FP16
              // storage and for the learned path's V-cycle step after a trial
              && ((cfg->anderson_m == 0 && cfg->precond_dtype != MGX_F16
                      && cfg->precond_kind != MGX_PRECOND_LEARNED)
                  || hipMalloc(&s->kr1, s->npts * tsize) == hipSuccess)
              && (cfg->precond_kind != MGX_PRECOND_LEARNED
                  || (alloc(&s->psi_prev, s->npts) && alloc(&s->r_prev, s->npts) && alloc(&s->psi_cnn, s->npts)
                      && alloc(&s->r_cnn, s->npts)
                      && hipMalloc(&s->kcnn, s->npts * tsize) == hipSuccess))
              && (cfg->lap_type != MGX_LAP_MEHRSTELLEN
                  || (alloc(&s->bpsi, s->npts) && alloc(&s->Bm, nn) && alloc(&s->invB, nn)))
              && (cfg->lap_type == MGX_LAP_MEHRSTELLEN || alloc(&s->xprev, nn))
              && hipHostMalloc((void**)&s->host, 32 * sizeof(double), hipHostMallocDefault) == hipSuccess;
    if (!ok)
    {
        free_solver(s);
        delete s;
        return set_err(ctx, MGX_ENOMEM, "solver allocation failed");
    }
    double vmax = vtot_host[0];
    for (int64_t i = 1; i < gsize; i++)
        vmax = std::max(vmax, vtot_host[i]);
    s->vmax = vmax;
    hipMemcpy(s->vtot, vtot_host, gsize * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(s->psi0, psi0_host, s->npts * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(s->psi, s->psi0, s->npts * sizeof(double), hipMemcpyDeviceToDevice);
    // the history and correction buffers start at zero (the kernels read the
    // history only once it exists; zero keeps a stray read finite)
    for (void* p : { (void*)s->x0, s->kr0, s->kr1, s->kcnn })
        if (p) hipMemset(p, 0, p == (void*)s->x0 ? s->npts * sizeof(double) : s->npts * tsize);
    if (hipDeviceSynchronize() != hipSuccess)
    {
        free_solver(s);
        delete s;
        return set_err(ctx, MGX_EHIP, "solver upload failed");
    }
    *out = s;
    return MGX_OK;
}

int mgx_solver_destroy(mgx_solver* s)
{
    if (!s) return MGX_OK;
    hipStreamSynchronize(s->ctx->stream);
    free_solver(s);
    delete s;
    return MGX_OK;
}

int mgx_solver_step(mgx_solver* s, mgx_solver_stats_t* st)
{
    if (!s || !st) return MGX_EINVAL;
    bool done = false;
    auto t0   = std::chrono::steady_clock::now();
    int rc    = step_impl(s, -1., st, &done);
    if (!rc) rc = flush_norm(s);
    st->seconds    = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st->iterations = s->iter;
    st->converged  = 0;
    return rc;
}

int mgx_solver_run(mgx_solver* s, double tol, int maxit, mgx_solver_stats_t* st)
{
    if (!s || !st || maxit < 1) return MGX_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    bool done = false;
    int rc    = MGX_OK;
    int it0   = s->iter;
    for (int k = 0; k < maxit && !done; k++)
    {
        rc = step_impl(s, tol, st, &done);
        if (rc) break;
    }
    if (!rc) rc = flush_norm(s);
    st->seconds    = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st->iterations = s->iter - it0;
    st->converged  = done ? 1 : 0;
    return rc;
}

int mgx_solver_reset(mgx_solver* s)
{
    if (!s) return MGX_EINVAL;
    MGX_CHECK_HIP(s->ctx, hipMemcpyAsync(s->psi, s->psi0, s->npts * sizeof(double), hipMemcpyDeviceToDevice,
                              s->ctx->stream));
    MGX_CHECK_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    s->iter       = 0;
    s->lowest     = 0.;
    s->have_inv   = false;
    s->have_xprev = false;
    s->pending_norm = false;
    s->e_prev     = 0.;
    s->deig       = 1e9;
    s->deig2      = 1e9;
    s->accelerate = false;
    s->mm         = 0;
    s->trial         = 0;
    s->backoff       = 0;
    s->backoff_len   = MGX_LEARNED_BACKOFF;
    s->learned_steps = 0;
    s->fallbacks     = 0;
    s->ns_steps      = 0;
    s->ns_pred       = 2;
    s->chol_calls    = 0;
    s->dots_sweeps   = 0;
    s->ns_continued  = 0;
    s->ns_restarts   = 0;
    s->last_more     = 0;
    return MGX_OK;
}

int mgx_solver_profile(mgx_solver* s, double* ms, int64_t* launches, int reset)
{
    if (!s) return MGX_EINVAL;
    for (int g = 0; g < MGX_PROF_N; g++)
    {
        if (ms) ms[g] = s->prof.ms[g];
        if (launches) launches[g] = s->prof.cnt[g];
        if (reset)
        {
            s->prof.ms[g]   = 0.;
            s->prof.cnt[g]  = 0;
            s->prof.wsum[g] = 0.;
        }
    }
    return MGX_OK;
}

int mgx_solver_set_profile(mgx_solver* s, int on)
{
    if (!s) return MGX_EINVAL;
    s->cfg.profile = on ? 1 : 0;
    return MGX_OK;
}

int mgx_solver_profile_work(mgx_solver* s, double* work)
{
    if (!s || !work) return MGX_EINVAL;
    for (int g = 0; g < MGX_PROF_N; g++)
        work[g] = s->prof.wsum[g];
    return MGX_OK;
}

int mgx_solver_get_psi(mgx_solver* s, double* host)
{
    if (!s || !host) return MGX_EINVAL;
    int rc = flush_norm(s);
    if (rc) return rc;
    MGX_CHECK_HIP(s->ctx, hipMemcpy(host, s->psi, s->npts * sizeof(double), hipMemcpyDeviceToHost));
    return MGX_OK;
}

int mgx_solver_inverse(mgx_solver* s, double** Sinv)
{
    if (!s || !Sinv) return MGX_EINVAL;
    *Sinv = s->Sinv;
    return MGX_OK;
}

int mgx_solver_set_psi(mgx_solver* s, const double* host)
{
    if (!s || !host) return MGX_EINVAL;
    MGX_CHECK_HIP(s->ctx, hipMemcpyAsync(s->psi, host, s->npts * sizeof(double), hipMemcpyHostToDevice,
                              s->ctx->stream));
    MGX_CHECK_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    // the caller's orbitals as they are: no pending normalisation, no
    // Anderson history, no network trial in flight; S^-1 and the step
    // counts stay as the refresh's start
    s->pending_norm = false;
    s->mm           = 0;
    s->trial        = 0;
    s->have_xprev   = false;  // the orbitals jumped: no extrapolation from the last two inverses
    return MGX_OK;
}

int mgx_solver_set_keep_residual(mgx_solver* s, int on)
{
    if (!s) return MGX_EINVAL;
    s->keep_res = on != 0;
    return MGX_OK;
}

int mgx_solver_set_graph(mgx_solver* s, int on)
{
    if (!s) return MGX_EINVAL;
    s->use_graph = on != 0;
    return MGX_OK;
}

int mgx_solver_graph_stats(mgx_solver* s, int64_t* launches, int64_t* captures, int64_t* cached)
{
    if (!s) return MGX_EINVAL;
    if (launches) *launches = s->graph_launches;
    if (captures) *captures = s->graph_captures;
    if (cached)
    {
        int64_t c = 0;
        for (auto& kv : s->graphs)
            c += kv.second.exec ? 1 : 0;
        *cached = c;
    }
    return MGX_OK;
}

int mgx_solver_set_cnn(mgx_solver* s, mgx_cnn* cnn)
{
    if (!s) return MGX_EINVAL;
    s->cnn = cnn;
    return MGX_OK;
}

int mgx_solver_density(mgx_solver* s, double nel, double* rho, double* charge)
{
    if (!s) return MGX_EINVAL;
    MGX_REQUIRE(s->ctx, rho, "null rho");
    mgx_ctx* ctx = s->ctx;
    int rc       = flush_norm(s);
    if (rc) return rc;
    // S of the current psi (FP64; H is a by-product here), its inverse, and
    // dm = 2 S^-1 in theta's buffer
    if ((rc = box_gram_h(ctx, s->bs, MGX_PREC_F64, s->psi, s->psi, s->S, s->H, s->Sv, s->Hv))) return rc;
    if ((rc = chol_inv(ctx, s->n, s->S, s->L, s->Sinv))) return rc;
    s->have_inv      = true;
    s->have_xprev    = false;
    const int64_t nn = (int64_t)s->n * s->n;
    hipLaunchKernelGGL(k_scaled_copy, dim3(nblocks(nn)), dim3(kBlock), 0, ctx->stream, (const double*)s->Sinv, nn, 2.,
        s->th);
    MGX_CHECK_LAUNCH(ctx);
    if ((rc = box_density(ctx, s->bs, s->psi, s->th, s->n, nel, rho, charge))) return rc;
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MGX_OK;
}

int mgx_solver_buffers(mgx_solver* s, double** psi, double** hpsi, double** S, double** H, double** res)
{
    if (!s) return MGX_EINVAL;
    int rc = flush_norm(s);
    if (rc) return rc;
    if (psi) *psi = s->psi;
    if (hpsi) *hpsi = s->hpsi;
    if (S) *S = s->S;
    if (H) *H = s->H;
    if (res) *res = s->res;
    return MGX_OK;
}

}  // extern "C"
