// C ABI implementation (include/ccoil_hip.h): context, device tables, buffers, field entry point.
#include "context.hpp"

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstring>

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
    for (int64_t i = 0; i < c->n; ++i) {
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
        out[(size_t)i] = d;
    }
    return "";
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

// Launch-shape rule of a field call. Split the thickness rows over 4 lanes when there are too few points to give
// every SIMD of the 256 CUs several waves, and over 16 lanes when even one chunk per coil would leave the launch
// below one wave per SIMD (few points and few coils: then each lane's serial chain of nodes, not the issue rate,
// sets the kernel's duration; one coil at 4,096 points measured 46 us at 4 lanes and 25 us at 16, while at 65,536
// points 16 lanes were slower). Split the coils into chunks (grid y, summed in order by field_reduce_kernel) while
// the launch still has fewer than about 8 waves per SIMD. Large launches take 1-4 chunks, whichever makes the last
// round of resident waves fullest: every wave does the same work, so with R rounds of resident waves the launch
// runs ceil(R) rounds.
FieldShape field_shape_rule(ccoil_ctx ctx, int64_t n, int64_t n_coils, uint32_t q, int precision)
{
    FieldShape sh;
    const int64_t target = (int64_t)256 * 4 * 8;
    sh.split = n < (int64_t)256 * 4 * 64 * 4 ? 4 : 1;
    if (sh.split == 4 && (n * 4 + 63) / 64 * n_coils < (int64_t)256 * 4) sh.split = 16;
    const int64_t waves = (n * sh.split + 63) / 64;
    int64_t chunks = waves >= target ? 1 : std::min<int64_t>(n_coils, (target + waves - 1) / waves);
    if (waves >= target && n * 18 < ((int64_t)1 << 31)) {   // the reduction grid stays below 2^31 threads
        const int bpc = field_blocks_per_cu(q, precision, sh.split);
        const double slots = (double)ctx->cu_count * bpc * (kFieldBlock / 64);
        if (slots > 0) {
            double best = 0.0;
            for (int64_t c = 1; c <= std::min<int64_t>(4, n_coils); ++c) {
                const double R = (double)waves * c / slots, eff = R / std::ceil(R);
                if (eff > best + 0.02) {
                    best = eff;
                    chunks = c;
                }
            }
        }
    }
    sh.chunks = (int)std::max<int64_t>(1, std::min<int64_t>(chunks, 65535));
    return sh;
}

// Sizes the near pass's buffers of L for cap records and points L.p at them (the context's grow-only buffers).
ccoil_status near_buffers(ccoil_ctx ctx, FieldLaunch &L, int64_t cap, int width)
{
    FieldParams &p = L.p;
    const int64_t n = p.n_pts;
    CC_HIP(ctx, ensure(ctx->ncounts, sizeof(int64_t) * (size_t)n));
    CC_HIP(ctx, ensure(ctx->noffsets, sizeof(int64_t) * (size_t)n));
    CC_HIP(ctx, ensure(ctx->nhist, 2 * sizeof(uint32_t) * ((size_t)p.n_coils + 2)));
    CC_HIP(ctx, ensure(ctx->nstate, sizeof(NearState)));
    size_t tb = 0;
    CC_HIP(ctx, near_scan(ctx->stream, nullptr, nullptr, n, nullptr, tb));
    CC_HIP(ctx, ensure(ctx->ntemp, tb > 0 ? tb : 1));
    cap = std::max<int64_t>(cap, 1);
    CC_HIP(ctx, ensure(ctx->nkeys, sizeof(uint32_t) * (size_t)cap));
    CC_HIP(ctx, ensure(ctx->nrecs, sizeof(uint64_t) * (size_t)cap));
    CC_HIP(ctx, ensure(ctx->nvals, sizeof(double) * (size_t)width * (size_t)cap));
    ctx->near_cap = std::max(ctx->near_cap, cap);
    p.near_fp64 = 1;
    p.near_counts = (int64_t *)ctx->ncounts.p;
    p.near_offsets = (const int64_t *)ctx->noffsets.p;
    p.near_hist = (uint32_t *)ctx->nhist.p;
    p.near_cursor = p.near_hist + ((size_t)p.n_coils + 2);
    p.near_state = (NearState *)ctx->nstate.p;
    p.near_cap = cap;
    p.near_keys = (uint32_t *)ctx->nkeys.p;
    p.near_recs = (uint64_t *)ctx->nrecs.p;
    p.near_vals = (double *)ctx->nvals.p;
    p.near_width = width;
    L.scan_temp = ctx->ntemp.p;
    L.scan_bytes = tb;
    return CCOIL_OK;
}

// Checks the near pass's outcome after the call's work is done.
ccoil_status near_outcome(ccoil_ctx ctx, const NearState &st)
{
    if (st.mismatch)
        return fail(ctx, CCOIL_ERR_DEVICE, "FP32C near pass: the fill classified a (coil, point) pair differently "
                                           "from the field kernel; the near contributions are incomplete");
    return CCOIL_OK;
}

void enqueue_near_tail(hipStream_t s, const FieldLaunch &L, NearState *host_state)
{
    launch_near_key_scan(s, L.p);
    launch_near_fill(s, L.p);
    launch_near_eval(s, L.p);
    launch_near_add(s, L.p);
    if (host_state) (void)hipMemcpyAsync(host_state, L.p.near_state, sizeof(NearState), hipMemcpyDeviceToHost, s);
}

void enqueue_fields(hipStream_t s, const FieldLaunch &L, hipEvent_t evm0, hipEvent_t evm1, NearState *host_state)
{
    const FieldParams &p = L.p;
    if (L.near) {
        launch_fill_zero(s, p.near_counts, p.n_pts);
        launch_fill_zero(s, p.near_hist, ((int64_t)p.n_coils + 2) / 2);
        launch_fill_zero(s, p.near_state, sizeof(NearState) / 8);
    }
    if (evm0) (void)hipEventRecord(evm0, s);
    launch_field(s, p, L.precision, L.split);
    if (evm1) (void)hipEventRecord(evm1, s);
    launch_field_reduce(s, p);
    if (L.near) {
        size_t tb = L.scan_bytes;
        (void)near_scan(s, p.near_counts, const_cast<int64_t *>(p.near_offsets), p.n_pts, L.scan_temp, tb);
        enqueue_near_tail(s, L, host_state);
    }
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
    if ((e = hipHostMalloc((void **)&ctx->near_host, sizeof(NearState), hipHostMallocDefault)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    const size_t nd = sizeof(double) * kTableSize, nf = sizeof(float) * kTableSize;
    double **dp[11] = {&ctx->gl_x, &ctx->gl_w, &ctx->gl_cos, &ctx->gl_omc, &ctx->gl_rcos, &ctx->gl_rsin, &ctx->gl_wcos,
                       &ctx->gl_opc, &ctx->gl_ws, &ctx->gl_ws2, &ctx->gl_wcs};
    float **fp[7] = {&ctx->gl_xf, &ctx->gl_wf, &ctx->gl_cosf, &ctx->gl_omcf, &ctx->gl_wcosf, &ctx->gl_opcf,
                     &ctx->gl_wcsf};
    for (auto p : dp)
        if ((e = hipMalloc((void **)p, nd)) != hipSuccess) return bail(e, "hipMalloc(tables)");
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
                      &ctx->nvals, &ctx->ntemp, &ctx->pprim, &ctx->psec, &ctx->ppairs, &ctx->ppartial, &ctx->pcounter, &ctx->pM, &ctx->pF, &ctx->pT, &ctx->pdM, &ctx->acoils, &ctx->apts, &ctx->agB,
                      &ctx->apartial, &ctx->agrad, &ctx->zitems, &ctx->zpartial, &ctx->zout};
    for (auto b : bufs) free_buf(*b);
    if (ctx->near_host) (void)hipHostFree(ctx->near_host);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->evm0) (void)hipEventDestroy(ctx->evm0);
    if (ctx->evm1) (void)hipEventDestroy(ctx->evm1);
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

ccoil_status ccoil_fields(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                          uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B, double *E,
                          double *G, ccoil_report *report)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_fields: ctx is NULL");
    if (n_points < 0) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_points is negative");
    if (n_points > ((int64_t)1 << 31) - 1) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_points exceeds 2^31 - 1");
    if (precision != CCOIL_FP64 && precision != CCOIL_FP32C && precision != CCOIL_FP32 &&
        precision != CCOIL_FP64_REFERENCE && precision != CCOIL_FP64_IEEE)
        return fail(ctx, CCOIL_ERR_ARGUMENT, "unknown precision");
    const uint32_t known = CCOIL_A | CCOIL_B | CCOIL_E | CCOIL_G | CCOIL_G_JACOBIAN;
    if (quantities == 0 || (quantities & ~known)) return fail(ctx, CCOIL_ERR_ARGUMENT, "bad quantity mask");
    if ((quantities & CCOIL_G) && (quantities & CCOIL_G_JACOBIAN))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "CCOIL_G and CCOIL_G_JACOBIAN share the G output; request one");
    if (((quantities & CCOIL_A) && !A) || ((quantities & CCOIL_B) && !B) || ((quantities & CCOIL_E) && !E) ||
        ((quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) && !G))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "an output for a requested quantity is NULL");
    if (n_points > 0 && !points) return fail(ctx, CCOIL_ERR_ARGUMENT, "points is NULL");
    std::vector<CoilDev> hc;
    std::string msg = pack_coils(coils, false, (quantities & CCOIL_E) != 0, hc);
    if (!msg.empty()) return fail(ctx, CCOIL_ERR_ARGUMENT, msg);
    if (hc.size() > (size_t)INT32_MAX) return fail(ctx, CCOIL_ERR_ARGUMENT, "too many coils");
    if (report) std::memset(report, 0, sizeof *report);
    CC_HIP(ctx, hipSetDevice(ctx->device));
    const bool dev = flags & CCOIL_DEVICE_POINTERS;
    const int64_t n = n_points;
    const size_t s3 = sizeof(double) * 3 * (size_t)n, s9 = sizeof(double) * 9 * (size_t)n;

    if (n == 0) return CCOIL_OK;
    if (hc.empty()) {
        // No coils: every field is zero (the reference returns a zero-filled array).
        double *outs[4] = {(quantities & CCOIL_A) ? A : nullptr, (quantities & CCOIL_B) ? B : nullptr,
                           (quantities & CCOIL_E) ? E : nullptr,
                           (quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) ? G : nullptr};
        size_t sz[4] = {s3, s3, s3, s9};
        for (int k = 0; k < 4; ++k) {
            if (!outs[k]) continue;
            if (dev) launch_fill_zero(ctx->stream, outs[k], (int64_t)(sz[k] / sizeof(double)));
            else std::memset(outs[k], 0, sz[k]);
        }
        if (dev) CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return CCOIL_OK;
    }

    // The FP32C near pass sorts its (coil, point) records by near_key: coils with the most nodes first, so its
    // longest waves start first. near_coil (near_key -> coil) follows the coils in the same buffer.
    std::vector<uint32_t> near_coil(hc.size());
    for (size_t i = 0; i < hc.size(); ++i) near_coil[i] = (uint32_t)i;
    auto cost = [&](uint32_t i) { return (int64_t)hc[i].n_ang * hc[i].n_thk; };   // field nodes of the coil
    std::stable_sort(near_coil.begin(), near_coil.end(), [&](uint32_t x, uint32_t y) { return cost(x) > cost(y); });
    for (size_t k = 0; k < hc.size(); ++k) hc[near_coil[k]].near_key = (int32_t)k;
    const size_t coil_bytes = sizeof(CoilDev) * hc.size();
    CC_HIP(ctx, ensure(ctx->coils, coil_bytes + sizeof(uint32_t) * hc.size()));
    CC_HIP(ctx, hipMemcpyAsync(ctx->coils.p, hc.data(), coil_bytes, hipMemcpyHostToDevice, ctx->stream));
    CC_HIP(ctx, hipMemcpyAsync((char *)ctx->coils.p + coil_bytes, near_coil.data(), sizeof(uint32_t) * hc.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    FieldLaunch L;
    FieldParams &p = L.p;
    std::memset(&p, 0, sizeof p);
    p.coils = (const CoilDev *)ctx->coils.p;
    p.n_coils = (int32_t)hc.size();
    p.near_coil = (const uint32_t *)((const char *)ctx->coils.p + coil_bytes);
    p.n_pts = n;
    p.q = quantities;
    p.tab = ctx_tables(ctx);
    if (dev) {
        p.pts = points;
        p.A = (quantities & CCOIL_A) ? A : nullptr;
        p.B = (quantities & CCOIL_B) ? B : nullptr;
        p.E = (quantities & CCOIL_E) ? E : nullptr;
        p.G = (quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) ? G : nullptr;
    } else {
        CC_HIP(ctx, ensure(ctx->pts, s3));
        CC_HIP(ctx, hipMemcpyAsync(ctx->pts.p, points, s3, hipMemcpyHostToDevice, ctx->stream));
        p.pts = (const double *)ctx->pts.p;
        if (quantities & CCOIL_A) { CC_HIP(ctx, ensure(ctx->outA, s3)); p.A = (double *)ctx->outA.p; }
        if (quantities & CCOIL_B) { CC_HIP(ctx, ensure(ctx->outB, s3)); p.B = (double *)ctx->outB.p; }
        if (quantities & CCOIL_E) { CC_HIP(ctx, ensure(ctx->outE, s3)); p.E = (double *)ctx->outE.p; }
        if (quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) { CC_HIP(ctx, ensure(ctx->outG, s9)); p.G = (double *)ctx->outG.p; }
    }
    const int width = (p.A ? 3 : 0) + (p.E ? 3 : 0) + (p.B ? 3 : 0) + (p.G ? 9 : 0);
    const FieldShape shape = field_shape_rule(ctx, n, (int64_t)hc.size(), quantities, precision);
    L.precision = precision;
    L.split = shape.split;
    p.coils_per_chunk = (int32_t)(((int64_t)hc.size() + shape.chunks - 1) / shape.chunks);
    p.n_chunks = (int32_t)(((int64_t)hc.size() + p.coils_per_chunk - 1) / p.coils_per_chunk);
    if (p.n_chunks > 1) {
        p.partial_width = width;
        CC_HIP(ctx, ensure(ctx->fpartial, sizeof(double) * p.partial_width * (size_t)n * (size_t)p.n_chunks));
        p.partial = (double *)ctx->fpartial.p;
    }
    L.near = precision == CCOIL_FP32C && !(flags & CCOIL_NO_NEAR_FP64);
    if (L.near) {
        // Buffers for the records: every (coil, point) pair when that is small, else a capacity kept from earlier
        // calls (at least two per point); a call that finds more grows them and runs the near pass again.
        const int64_t all_pairs = n * (int64_t)hc.size();
        int64_t cap = std::max<int64_t>(ctx->near_cap, std::min<int64_t>(all_pairs, 2 * n + 65536));
        cap = std::min<int64_t>(cap, INT32_MAX);
        ccoil_status st = near_buffers(ctx, L, cap, width);
        if (st != CCOIL_OK) return st;
    }
    CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    enqueue_fields(ctx->stream, L, ctx->evm0, ctx->evm1, L.near ? ctx->near_host : nullptr);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    auto copy_out = [&]() -> ccoil_status {
        if (dev) return CCOIL_OK;
        if (p.A) CC_HIP(ctx, hipMemcpyAsync(A, p.A, s3, hipMemcpyDeviceToHost, ctx->stream));
        if (p.B) CC_HIP(ctx, hipMemcpyAsync(B, p.B, s3, hipMemcpyDeviceToHost, ctx->stream));
        if (p.E) CC_HIP(ctx, hipMemcpyAsync(E, p.E, s3, hipMemcpyDeviceToHost, ctx->stream));
        if (p.G) CC_HIP(ctx, hipMemcpyAsync(G, p.G, s9, hipMemcpyDeviceToHost, ctx->stream));
        return CCOIL_OK;
    };
    ccoil_status st = copy_out();
    if (st != CCOIL_OK) return st;
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t near = 0;
    if (L.near) {
        st = near_outcome(ctx, *ctx->near_host);
        if (st != CCOIL_OK) return st;
        near = (int64_t)ctx->near_host->total;
        if (ctx->near_host->overflow) {
            // more near pairs than the buffers hold: the near pass did nothing; grow and run it again
            st = near_buffers(ctx, L, std::min<int64_t>(near + near / 8 + 4096, INT32_MAX), width);
            if (st != CCOIL_OK) return st;
            enqueue_near_tail(ctx->stream, L, ctx->near_host);
            CC_HIP(ctx, hipGetLastError());
            CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
            st = copy_out();
            if (st != CCOIL_OK) return st;
            CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            st = near_outcome(ctx, *ctx->near_host);
            if (st != CCOIL_OK) return st;
            if (ctx->near_host->overflow)
                return fail(ctx, CCOIL_ERR_DEVICE, "FP32C near pass: capacity still exceeded after growing");
        }
    }
    if (report) {
        float ms = 0.f, mms = 0.f;
        CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        CC_HIP(ctx, hipEventElapsedTime(&mms, ctx->evm0, ctx->evm1));
        int64_t nodes = 0;
        for (const auto &c : hc) nodes += (int64_t)c.n_ang * c.n_thk;
        report->node_evaluations = nodes * n;
        report->near_pairs = near;
        report->kernel_ms = ms;
        report->main_kernel_ms = mms;
        report->block_size = kFieldBlock;
        report->split = L.split;
    }
    return CCOIL_OK;
}

}  // extern "C"
