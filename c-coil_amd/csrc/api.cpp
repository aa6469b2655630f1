// C ABI implementation (include/ccoil_hip.h): context, device tables, buffers, field entry point.
#include "context.hpp"

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

using namespace ccoil;

namespace {
thread_local std::string g_create_error;
}

namespace ccoil {

ccoil_status fail(ccoil_ctx ctx, ccoil_status code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    else g_create_error = msg;
    return code;
}

hipError_t ensure(DevBuf &b, size_t bytes)
{
    if (bytes <= b.cap) return hipSuccess;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return e;
    }
    size_t want = bytes + bytes / 4 + 256;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return e;
    }
    b.cap = want;
    return hipSuccess;
}

void free_buf(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

ccoil_status call_buffer(ccoil_ctx ctx, DevBuf &b, Arena *arena, size_t bytes, void **out)
{
    if (arena) {
        *out = arena->take(bytes);
        if (!*out)
            return fail(ctx, CCOIL_ERR_ARGUMENT, "the call needs more device memory than its fixed scratch holds (" +
                                                 std::to_string(arena->cap) + " bytes); split it");
        return CCOIL_OK;
    }
    CC_HIP(ctx, ensure(b, bytes));
    *out = b.p;
    return CCOIL_OK;
}

// Transformation matrices of Coil::calculateTransformationMatrices (CalculateAttributes.cxx:63-78).
void rotation(double alpha, double beta, double *T, double *Ti)
{
    const double cosY = std::cos(alpha), sinY = std::sin(alpha);
    const double cosZ = std::cos(beta), sinZ = std::sin(beta);
    const double t[9] = {cosZ * cosZ * cosY - sinZ * sinZ, -sinZ * cosZ * cosY - sinZ * cosZ, cosZ * sinY,
                         sinZ * cosZ * cosY + sinZ * cosZ, cosZ * cosZ - sinZ * sinZ * cosY, sinZ * sinY,
                         -sinY * cosZ, sinY * sinZ, cosY};
    const double ti[9] = {cosZ * cosZ * cosY - sinZ * sinZ, sinZ * cosZ * cosY + sinZ * cosZ, -cosZ * sinY,
                          -sinZ * cosZ * cosY - sinZ * cosZ, cosZ * cosZ - sinZ * sinZ * cosY, sinZ * sinY,
                          sinY * cosZ, sinY * sinZ, cosY};
    std::memcpy(T, t, sizeof t);
    std::memcpy(Ti, ti, sizeof ti);
}

}  // namespace ccoil

namespace ccoil {

// Axial vectors of Omega = T^T dT/dtheta for theta = alpha, beta, with T of rotation() differentiated entry by
// entry (the antisymmetric part is read off; T^T dT is antisymmetric because T is orthogonal).
void rotation_generators(double alpha, double beta, double *om_a, double *om_b)
{
    const double cY = std::cos(alpha), sY = std::sin(alpha), cZ = std::cos(beta), sZ = std::sin(beta);
    double T[9], Ti[9];
    rotation(alpha, beta, T, Ti);
    const double dA[9] = {-cZ * cZ * sY, sZ * cZ * sY, cZ * cY,
                          -sZ * cZ * sY, sZ * sZ * sY, sZ * cY,
                          -cY * cZ, cY * sZ, -sY};
    const double dB[9] = {-2 * cZ * sZ * cY - 2 * sZ * cZ, -(cZ * cZ - sZ * sZ) * cY - (cZ * cZ - sZ * sZ), -sZ * sY,
                          (cZ * cZ - sZ * sZ) * cY + (cZ * cZ - sZ * sZ), -2 * cZ * sZ - 2 * sZ * cZ * cY, cZ * sY,
                          sY * sZ, sY * cZ, 0.0};
    double *om[2] = {om_a, om_b};
    const double *dT[2] = {dA, dB};
    for (int k = 0; k < 2; ++k) {
        double O[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                O[3 * i + j] = T[i] * dT[k][j] + T[3 + i] * dT[k][3 + j] + T[6 + i] * dT[k][6 + j];
        om[k][0] = 0.5 * (O[7] - O[5]);
        om[k][1] = 0.5 * (O[2] - O[6]);
        om[k][2] = 0.5 * (O[3] - O[1]);
    }
}

// Validates a coil descriptor and packs coil i into the device layout. constFactor follows
// generateCoilData (SupportFunctions.cxx:169-172): 1e-7 J a pi/2, times b for the slow method.
std::string pack_coils(const ccoil_coils *c, bool need_len, bool need_freq, std::vector<CoilDev> &out)
{
    if (!c) return "coil descriptor is NULL";
    if (c->n < 0) return "coil count is negative";
    if (c->n == 0) {
        out.clear();
        return "";
    }
    if (!c->use_fast || !c->n_ang || !c->n_thk || !c->R || !c->a || !c->b || !c->J || !c->pos || !c->alpha ||
        !c->beta)
        return "coil descriptor has a NULL required array";
    if (need_len && (!c->n_len || !c->current || !c->turns)) return "pair coils need n_len, current and turns";
    if (need_freq && !c->frequency) return "E requested but coil frequency array is NULL";
    out.resize((size_t)c->n);
    // one coil's record, or the reason it has none
    auto pack_one = [&](int64_t i) -> std::string {
        CoilDev d;
        std::memset(&d, 0, sizeof d);
        const int na = c->n_ang[i], nt = c->n_thk[i], nl = c->n_len ? c->n_len[i] : 1;
        if (na < 1 || na > kMaxOrder || nt < 1 || nt > kMaxOrder || nl < 1 || nl > kMaxOrder)
            return "coil " + std::to_string(i) + ": increments must be in [1, 100]";
        if (!(c->R[i] > 0.0) || !(c->a[i] > 0.0) || !(c->b[i] > 0.0) || !std::isfinite(c->J[i]))
            return "coil " + std::to_string(i) + ": R, a, b must be positive and J finite";
        d.use_fast = c->use_fast[i] ? 1 : 0;
        d.n_ang = na;
        d.n_thk = nt;
        d.n_len = nl;
        d.blk_ang = c->blocks ? c->blocks[3 * i] : 1;
        d.blk_thk = c->blocks ? c->blocks[3 * i + 1] : 1;
        d.blk_len = c->blocks ? c->blocks[3 * i + 2] : 1;
        if (d.blk_ang < 1 || d.blk_thk < 1 || d.blk_len < 1) return "coil " + std::to_string(i) + ": blocks < 1";
        d.R = c->R[i];
        d.a = c->a[i];
        d.b = c->b[i];
        const double pi = kPi;
        d.cf = d.use_fast ? kMuReduced * c->J[i] * d.a * pi * 0.5 : kMuReduced * c->J[i] * d.a * d.b * pi * 0.5;
        d.omega = c->frequency ? 2.0 * pi * c->frequency[i] : 0.0;
        d.current = c->current ? c->current[i] : 0.0;
        d.turns = c->turns ? c->turns[i] : 0.0;
        double w = std::fmax(std::fmax(d.a, d.b), 1e-3 * d.R);
        d.near_w2 = w * w;
        for (int k = 0; k < 3; ++k) d.pos[k] = c->pos[3 * i + k];
        rotation(c->alpha[i], c->beta[i], d.T, d.Tinv);
        d.sa = std::sin(c->alpha[i]);
        d.ca = std::cos(c->alpha[i]);
        d.sb = std::sin(c->beta[i]);
        d.cb = std::cos(c->beta[i]);
        rotation_generators(c->alpha[i], c->beta[i], d.om_a, d.om_b);
        d.cur_used = d.current;
        d.adj_scale = 1.0;
        d.ni = c->J[i] * d.a * d.b;
        d.rmax = std::hypot(d.R + d.a, d.use_fast ? 0.5 * d.b : 0.0);   // the slow rule's winding is a flat sheet
        d.inv_rmax2 = 1.0 / (d.rmax * d.rmax);
        d.far_d2 = 0.0;
        d.mp_terms = 0;
        out[(size_t)i] = d;
        return "";
    };
    // ranges of coils on host threads; the error reported is the lowest coil's, as a serial loop would report
    std::vector<std::pair<int64_t, std::string>> errs;
    std::mutex m;
    host_parallel_for(c->n, [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            std::string e = pack_one(i);
            if (!e.empty()) {
                std::lock_guard<std::mutex> lock(m);
                errs.emplace_back(i, std::move(e));
                return;
            }
        }
    });
    if (errs.empty()) return "";
    return std::min_element(errs.begin(), errs.end())->second;
}

void host_parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &fn)
{
    constexpr int64_t kPerThread = 8192;
    const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
    const int64_t parts = std::min<int64_t>(std::min<int64_t>(16, hw), (n + kPerThread - 1) / kPerThread);
    if (parts <= 1) {
        if (n > 0) fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t per = (n + parts - 1) / parts;
    for (int64_t k = 1; k < parts; ++k) {
        const int64_t i0 = k * per, i1 = std::min(n, i0 + per);
        if (i0 < i1) pool.emplace_back(fn, i0, i1);
    }
    fn(0, std::min(n, per));
    for (auto &t : pool) t.join();
}

Tables ctx_tables(ccoil_ctx ctx)
{
    Tables t;
    t.x = ctx->gl_x;
    t.w = ctx->gl_w;
    t.cosv = ctx->gl_cos;
    t.omc = ctx->gl_omc;
    t.xf = ctx->gl_xf;
    t.wf = ctx->gl_wf;
    t.cosf = ctx->gl_cosf;
    t.omcf = ctx->gl_omcf;
    t.rcos = ctx->gl_rcos;
    t.wcos = ctx->gl_wcos;
    t.wcosf = ctx->gl_wcosf;
    t.opc = ctx->gl_opc;
    t.opcf = ctx->gl_opcf;
    t.ws = ctx->gl_ws;
    t.ws2 = ctx->gl_ws2;
    t.wcs = ctx->gl_wcs;
    t.wcsf = ctx->gl_wcsf;
    t.rsin = ctx->gl_rsin;
    return t;
}

}  // namespace ccoil

extern "C" {

const char *ccoil_version(void) { return "ccoil-mi355x 0.1.0 (gfx950)"; }

const char *ccoil_last_error(ccoil_ctx ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

ccoil_status ccoil_create(int device, ccoil_ctx *out)
{
    if (!out) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_create: out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(nullptr, CCOIL_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(nullptr, CCOIL_ERR_ARGUMENT, "device index out of range");
    ccoil_ctx ctx = new ccoil_ctx_s();
    ctx->device = device;
    auto bail = [&](hipError_t e, const char *what) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(e);
        ccoil_destroy(ctx);
        return e == hipErrorOutOfMemory ? CCOIL_ERR_OUT_OF_MEMORY : CCOIL_ERR_DEVICE;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    if ((e = hipDeviceGetAttribute(&ctx->cu_count, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return bail(e, "hipDeviceGetAttribute");
    if ((e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    ctx->stream = ctx->own_stream;
    if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ctx->ev1)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ctx->evm0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ctx->evm1)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ctx->evr0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ctx->evr1)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipHostMalloc((void **)&ctx->near_host, sizeof(NearState), hipHostMallocDefault)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    if ((e = scratch_address(device, &ctx->scratch)) != hipSuccess) return bail(e, "hipGetSymbolAddress(scratch)");
    if (const char *tf = std::getenv("CCOIL_TUNING_FILE")) load_tuning_file(ctx, tf);
    const size_t nd = sizeof(double) * kTableSize, nf = sizeof(float) * kTableSize;
    double **dp[11] = {&ctx->gl_x, &ctx->gl_w, &ctx->gl_cos, &ctx->gl_omc, &ctx->gl_rcos, &ctx->gl_rsin, &ctx->gl_wcos,
                       &ctx->gl_opc, &ctx->gl_ws, &ctx->gl_ws2, &ctx->gl_wcs};
    float **fp[7] = {&ctx->gl_xf, &ctx->gl_wf, &ctx->gl_cosf, &ctx->gl_omcf, &ctx->gl_wcosf, &ctx->gl_opcf,
                     &ctx->gl_wcsf};
    for (auto p : dp)   // x and w also hold the series' length rule (kTableSizeXW)
        if ((e = hipMalloc((void **)p, p == &ctx->gl_x || p == &ctx->gl_w ? sizeof(double) * kTableSizeXW : nd)) !=
            hipSuccess)
            return bail(e, "hipMalloc(tables)");
    for (auto p : fp)
        if ((e = hipMalloc((void **)p, nf)) != hipSuccess) return bail(e, "hipMalloc(tables)");
    launch_gl_tables(ctx->stream, ctx->gl_x, ctx->gl_w, ctx->gl_cos, ctx->gl_omc, ctx->gl_xf, ctx->gl_wf,
                     ctx->gl_cosf, ctx->gl_omcf, ctx->gl_rcos, ctx->gl_rsin);
    if ((e = hipGetLastError()) != hipSuccess) return bail(e, "gl_table_kernel launch");
    launch_angular_tables(ctx->stream, ctx->gl_x, ctx->gl_w, ctx->gl_cos, ctx->gl_omc, ctx->gl_wcos, ctx->gl_wcosf,
                          ctx->gl_opc, ctx->gl_opcf, ctx->gl_ws, ctx->gl_ws2, ctx->gl_wcs, ctx->gl_wcsf);
    if ((e = hipGetLastError()) != hipSuccess) return bail(e, "angular_table_kernel launch");
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return bail(e, "gl_table_kernel");
    *out = ctx;
    return CCOIL_OK;
}

ccoil_status ccoil_destroy(ccoil_ctx ctx)
{
    if (!ctx) return CCOIL_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
    double *dp[11] = {ctx->gl_x, ctx->gl_w, ctx->gl_cos, ctx->gl_omc, ctx->gl_rcos, ctx->gl_rsin, ctx->gl_wcos,
                      ctx->gl_opc, ctx->gl_ws, ctx->gl_ws2, ctx->gl_wcs};
    float *fp[7] = {ctx->gl_xf, ctx->gl_wf, ctx->gl_cosf, ctx->gl_omcf, ctx->gl_wcosf, ctx->gl_opcf, ctx->gl_wcsf};
    for (auto p : dp)
        if (p) (void)hipFree(p);
    for (auto p : fp)
        if (p) (void)hipFree(p);
    DevBuf *bufs[] = {&ctx->coils, &ctx->fpartial, &ctx->pts, &ctx->outA, &ctx->outB, &ctx->outE, &ctx->outG,
                      &ctx->ncounts, &ctx->noffsets, &ctx->nhist, &ctx->nstate, &ctx->nkeys, &ctx->nrecs,
                      &ctx->nvals, &ctx->ntemp, &ctx->fmp, &ctx->forder, &ctx->ftemp, &ctx->pprim, &ctx->psec, &ctx->ppairs, &ctx->ppartial, &ctx->pcounter, &ctx->pM, &ctx->pF, &ctx->pT, &ctx->pdM, &ctx->acoils, &ctx->apts, &ctx->agB,
                      &ctx->apartial, &ctx->agrad, &ctx->zitems, &ctx->zpartial, &ctx->zout};
    for (auto b : bufs) free_buf(*b);
    if (ctx->near_host) (void)hipHostFree(ctx->near_host);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->evm0) (void)hipEventDestroy(ctx->evm0);
    if (ctx->evm1) (void)hipEventDestroy(ctx->evm1);
    if (ctx->evr0) (void)hipEventDestroy(ctx->evr0);
    if (ctx->evr1) (void)hipEventDestroy(ctx->evr1);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return CCOIL_OK;
}

ccoil_status ccoil_set_stream(ccoil_ctx ctx, void *stream)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_set_stream: ctx is NULL");
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    return CCOIL_OK;
}

ccoil_status ccoil_get_stream(ccoil_ctx ctx, void **stream)
{
    if (!ctx || !stream) return fail(ctx, CCOIL_ERR_ARGUMENT, "ccoil_get_stream: ctx and stream must not be NULL");
    *stream = (void *)ctx->stream;
    return CCOIL_OK;
}

ccoil_status ccoil_synchronize(ccoil_ctx ctx)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_synchronize: ctx is NULL");
    CC_HIP(ctx, hipSetDevice(ctx->device));
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CCOIL_OK;
}

ccoil_status ccoil_gauss_legendre(ccoil_ctx ctx, int32_t n, double *x, double *w)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_gauss_legendre: ctx is NULL");
    if (n < 1 || n > kMaxOrder) return fail(ctx, CCOIL_ERR_ARGUMENT, "order must be in [1, 100]");
    if (!x || !w) return fail(ctx, CCOIL_ERR_ARGUMENT, "x and w must not be NULL");
    CC_HIP(ctx, hipSetDevice(ctx->device));
    const int off = table_offset(n);
    CC_HIP(ctx, hipMemcpyAsync(x, ctx->gl_x + off, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    CC_HIP(ctx, hipMemcpyAsync(w, ctx->gl_w + off, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CCOIL_OK;
}


}  // extern "C"
