// ctx.hip -- context lifetime, workspaces, error text.
#include "common.hpp"

namespace mgx
{

int set_err(mgx_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

void* workspace(mgx_ctx* ctx, int slot, size_t bytes)
{
    if (slot < 0 || slot >= mgx_ctx::kSlots) return nullptr;
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    if (ctx->ws_sz[slot] < bytes)
    {
        if (ctx->ws[slot])
        {
            hipStreamSynchronize(ctx->stream);
            hipFree(ctx->ws[slot]);
            ctx->ws[slot]    = nullptr;
            ctx->ws_sz[slot] = 0;
        }
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        ctx->ws[slot]    = p;
        ctx->ws_sz[slot] = bytes;
        ctx->ws_gen++;
    }
    return ctx->ws[slot];
}

int sync(mgx_ctx* ctx)
{
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MGX_OK;
}

int prof_begin(mgx_ctx* ctx, int group, double work)
{
    MgxProf* p = ctx->prof;
    if (!p || group < 0 || group >= kProfGroups) return -1;
    while (p->pool.size() < p->used + 2)
    {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        p->pool.push_back(e);
    }
    const int tok = (int)(p->used / 2);
    hipEventRecord(p->pool[p->used], ctx->stream);
    p->used += 2;
    if ((int)p->group.size() <= tok)
    {
        p->group.resize(tok + 1);
        p->work.resize(tok + 1);
    }
    p->group[tok] = group;
    p->work[tok]  = work;
    return tok;
}

void prof_end(mgx_ctx* ctx, int tok)
{
    if (!ctx->prof || tok < 0) return;
    hipEventRecord(ctx->prof->pool[2 * tok + 1], ctx->stream);
}

void prof_retag(mgx_ctx* ctx, int tok, int group, double work)
{
    MgxProf* p = ctx->prof;
    if (!p || tok < 0 || tok >= (int)p->group.size() || group < 0 || group >= kProfGroups) return;
    p->group[tok] = group;
    p->work[tok]  = work;
}

void prof_collect(mgx_ctx* ctx)
{
    MgxProf* p = ctx->prof;
    if (!p || p->used == 0) return;
    hipStreamSynchronize(ctx->stream);
    for (size_t t = 0; t < p->used / 2; t++)
    {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p->pool[2 * t], p->pool[2 * t + 1]) == hipSuccess)
        {
            p->ms[p->group[t]] += ms;
            p->cnt[p->group[t]] += 1;
            p->wsum[p->group[t]] += p->work[t];
        }
    }
    p->used = 0;
}

}  // namespace mgx

extern "C" {

int mgx_ctx_create(int device, mgx_ctx** out)
{
    if (!out) return MGX_EINVAL;
    *out       = nullptr;
    int ndev   = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MGX_EHIP;
    if (device < 0 || device >= ndev) return MGX_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MGX_EHIP;
    mgx_ctx* c = new mgx_ctx();
    c->device  = device;
    // default priority stream
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    {
        delete c;
        return MGX_EHIP;
    }
    if (rocblas_create_handle(&c->blas) != rocblas_status_success)
    {
        hipStreamDestroy(c->stream);
        delete c;
        return MGX_EHIP;
    }
    rocblas_set_stream(c->blas, c->stream);
    if (hipHostMalloc((void**)&c->host_pinned, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        c->host_pinned = nullptr;
    // the device address of this GPU's no-allocation scratch array
    if (mgx::scratch_init(c, device) != MGX_OK)
    {
        mgx_ctx_destroy(c);
        return MGX_EHIP;
    }
    *out = c;
    return MGX_OK;
}

int mgx_ctx_destroy(mgx_ctx* c)
{
    if (!c) return MGX_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (int i = 0; i < mgx_ctx::kSlots; i++)
        if (c->ws[i]) hipFree(c->ws[i]);
    if (c->host_pinned) hipHostFree(c->host_pinned);
    if (c->blas) rocblas_destroy_handle(c->blas);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return MGX_OK;
}

const char* mgx_last_error(const mgx_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* mgx_ctx_stream(mgx_ctx* c) { return c ? (void*)c->stream : nullptr; }

// MemorySpace::Memory<T, MemorySpace::Device>::allocate / free and
// MemorySpace::copy_to_dev / copy_to_host, src/tools/memory_space.h
int mgx_malloc(mgx_ctx* c, size_t bytes, void** ptr)
{
    if (!c || !ptr) return MGX_EINVAL;
    *ptr = nullptr;
    if (bytes == 0) return MGX_OK;
    hipSetDevice(c->device);
    if (hipMalloc(ptr, bytes) != hipSuccess) return mgx::set_err(c, MGX_ENOMEM, "hipMalloc failed");
    return MGX_OK;
}

int mgx_free(mgx_ctx* c, void* ptr)
{
    if (!c) return MGX_EINVAL;
    if (!ptr) return MGX_OK;
    hipStreamSynchronize(c->stream);
    MGX_CHECK_HIP(c, hipFree(ptr));
    return MGX_OK;
}

int mgx_memcpy_htod(mgx_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (!c || (bytes && (!dst || !src))) return MGX_EINVAL;
    if (!bytes) return MGX_OK;
    MGX_CHECK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return mgx::sync(c);
}

int mgx_memcpy_dtoh(mgx_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (!c || (bytes && (!dst || !src))) return MGX_EINVAL;
    if (!bytes) return MGX_OK;
    MGX_CHECK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return mgx::sync(c);
}

double mgx_gamma(const double h[3], int mg_levels, double vmax, double small_eig)
{
    // OrbitalsPreconditioning<T>::setGamma, src/OrbitalsPreconditioning.cc:120-145
    double g = 1. / mgx::diag_el(MGX_LAP_4TH, h);
    for (int ln = 0; ln <= mg_levels; ln++)
        g *= 4.;
    return 1.0 / (2.0 / g + fabs(vmax - small_eig));
}

}  // extern "C"
