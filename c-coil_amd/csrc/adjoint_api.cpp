// C ABI: ccoil_field_gradient (adjoint of the summed B field, include/ccoil_hip.h).
#include "context.hpp"

#include <algorithm>
#include <cstring>

using namespace ccoil;

extern "C" ccoil_status ccoil_field_gradient(ccoil_ctx ctx, const ccoil_coils *coils, const double *points,
                                             int64_t n_points, const double *dL_dB, int32_t precision,
                                             uint32_t flags, double *grad, ccoil_report *report)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_field_gradient: ctx is NULL");
    if (n_points < 0) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_points is negative");
    if (precision != CCOIL_FP64 && precision != CCOIL_FP32C && precision != CCOIL_FP64_IEEE)
        return fail(ctx, CCOIL_ERR_ARGUMENT, "the adjoint supports CCOIL_FP64, CCOIL_FP32C and CCOIL_FP64_IEEE");
    if (!grad) return fail(ctx, CCOIL_ERR_ARGUMENT, "grad is NULL");
    if (n_points > 0 && (!points || !dL_dB)) return fail(ctx, CCOIL_ERR_ARGUMENT, "points or dL_dB is NULL");
    std::vector<CoilDev> hc;
    std::string msg = pack_coils(coils, false, false, hc);
    if (!msg.empty()) return fail(ctx, CCOIL_ERR_ARGUMENT, msg);
    if (!hc.empty() && (!coils->current || !coils->turns))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "the adjoint needs coil current and turns");
    if (hc.size() > 65535) return fail(ctx, CCOIL_ERR_ARGUMENT, "the adjoint takes at most 65535 coils per call");
    // B is linear in the current (J = N I / (a b)); a zero-current coil is evaluated at unit current so that
    // dL/dI is defined, and its other derivatives are scaled by I / 1 = 0.
    for (auto &d : hc) {
        if (d.current == 0.0) {
            const double J1 = d.turns / (d.a * d.b);
            d.cf = d.use_fast ? kMuReduced * J1 * d.a * kPi * 0.5 : kMuReduced * J1 * d.a * d.b * kPi * 0.5;
            d.cur_used = 1.0;
            d.adj_scale = 0.0;
        }
    }
    if (report) std::memset(report, 0, sizeof *report);
    CC_HIP(ctx, hipSetDevice(ctx->device));
    const bool dev = flags & CCOIL_DEVICE_POINTERS;
    const size_t gbytes = sizeof(double) * kGradTerms * hc.size();
    if (hc.empty()) return CCOIL_OK;
    if (n_points == 0) {
        if (dev) {
            launch_fill_zero(ctx->stream, grad, (int64_t)(gbytes / sizeof(double)));
            CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        } else {
            std::memset(grad, 0, gbytes);
        }
        return CCOIL_OK;
    }
    const size_t s3 = sizeof(double) * 3 * (size_t)n_points;
    CC_HIP(ctx, ensure(ctx->acoils, sizeof(CoilDev) * hc.size()));
    CC_HIP(ctx, hipMemcpyAsync(ctx->acoils.p, hc.data(), sizeof(CoilDev) * hc.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    GradParams p;
    std::memset(&p, 0, sizeof p);
    p.coils = (const CoilDev *)ctx->acoils.p;
    p.n_coils = (int32_t)hc.size();
    p.n_pts = n_points;
    p.tab = ctx_tables(ctx);
    // Point blocks: enough to give the CUs several waves over (blocks x coils), capped so that the partial-sum
    // buffer stays at blocks x coils x 9 doubles.
    const int64_t want = (n_points + 255) / 256;
    const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(4096, (int64_t)(1 << 22) / (int64_t)hc.size()));
    p.n_blocks = (int32_t)std::min(want, cap);
    CC_HIP(ctx, ensure(ctx->apartial, sizeof(double) * kGradTerms * hc.size() * (size_t)p.n_blocks));
    p.partial = (double *)ctx->apartial.p;
    if (dev) {
        p.pts = points;
        p.gB = dL_dB;
        p.grad = grad;
    } else {
        CC_HIP(ctx, ensure(ctx->apts, s3));
        CC_HIP(ctx, ensure(ctx->agB, s3));
        CC_HIP(ctx, ensure(ctx->agrad, gbytes));
        CC_HIP(ctx, hipMemcpyAsync(ctx->apts.p, points, s3, hipMemcpyHostToDevice, ctx->stream));
        CC_HIP(ctx, hipMemcpyAsync(ctx->agB.p, dL_dB, s3, hipMemcpyHostToDevice, ctx->stream));
        p.pts = (const double *)ctx->apts.p;
        p.gB = (const double *)ctx->agB.p;
        p.grad = (double *)ctx->agrad.p;
    }
    CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    launch_field_grad(ctx->stream, p, precision);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!dev) CC_HIP(ctx, hipMemcpyAsync(grad, p.grad, gbytes, hipMemcpyDeviceToHost, ctx->stream));
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (report) {
        float ms = 0.f;
        CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        int64_t nodes = 0;
        for (const auto &c : hc) nodes += (int64_t)c.n_ang * c.n_thk;
        report->node_evaluations = nodes * n_points;
        report->kernel_ms = ms;
        report->main_kernel_ms = ms;
        report->block_size = 256;
        report->split = 1;
    }
    return CCOIL_OK;
}
