// C ABI: ccoil_pairs_zaxis and ccoil_self_inductance (coaxial rules, include/ccoil_hip.h).
#include "context.hpp"
#include "zaxis.hpp"

#include <cmath>
#include <cstring>

using namespace ccoil;

namespace {

constexpr double kMu = 1e-7;   // g_MiReduced (CalculateMInductanceZAxis.cxx:12)

bool valid_rule(int32_t blocks, int32_t inc) { return blocks >= 1 && inc >= 1 && inc <= kMaxOrder; }

// Upload the items, run the node kernel and the ordered reduction, copy n_out results back.
ccoil_status run_items(ccoil_ctx ctx, std::vector<ZItemDev> &items, size_t n_out, uint32_t flags, double *out_host,
                       double *out_dev, int64_t node_evals, ccoil_report *report)
{
    int64_t wg = 0;
    for (auto &it : items) {
        it.first_wg = wg;
        wg += (it.units + kZBlock - 1) / kZBlock;
    }
    if (wg > INT32_MAX) return fail(ctx, CCOIL_ERR_ARGUMENT, "coaxial batch needs more than 2^31 - 1 workgroups");
    CC_HIP(ctx, hipSetDevice(ctx->device));
    CC_HIP(ctx, ensure(ctx->zitems, sizeof(ZItemDev) * items.size()));
    CC_HIP(ctx, ensure(ctx->zpartial, sizeof(double) * (size_t)wg));
    CC_HIP(ctx, hipMemcpyAsync(ctx->zitems.p, items.data(), sizeof(ZItemDev) * items.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    ZParams zp;
    std::memset(&zp, 0, sizeof zp);
    zp.items = (const ZItemDev *)ctx->zitems.p;
    zp.n_items = (int32_t)items.size();
    zp.n_wg = wg;
    zp.partial = (double *)ctx->zpartial.p;
    zp.tab = ctx_tables(ctx);
    const bool dev = flags & CCOIL_DEVICE_POINTERS;
    if (dev) {
        zp.out = out_dev;
    } else {
        CC_HIP(ctx, ensure(ctx->zout, sizeof(double) * n_out));
        zp.out = (double *)ctx->zout.p;
    }
    CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    launch_zaxis(ctx->stream, zp);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->evm1, ctx->stream));
    launch_zaxis_reduce(ctx->stream, zp);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!dev) CC_HIP(ctx, hipMemcpyAsync(out_host, zp.out, sizeof(double) * n_out, hipMemcpyDeviceToHost, ctx->stream));
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (report) {
        float ms = 0.f, mms = 0.f;
        CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        CC_HIP(ctx, hipEventElapsedTime(&mms, ctx->ev0, ctx->evm1));
        report->node_evaluations = node_evals;
        report->kernel_ms = ms;
        report->main_kernel_ms = mms;
        report->block_size = kZBlock;
        report->split = 1;
    }
    return CCOIL_OK;
}

std::string check_coils(const ccoil_coils *c, const char *who)
{
    if (!c) return std::string(who) + " descriptor is NULL";
    if (c->n < 0) return std::string(who) + " count is negative";
    if (c->n == 0) return "";
    if (!c->n_ang || !c->n_thk || !c->R || !c->a || !c->b || !c->J || !c->pos || !c->current || !c->turns)
        return std::string(who) + " descriptor needs n_ang, n_thk, R, a, b, J, pos, current and turns";
    for (int64_t i = 0; i < c->n; ++i) {
        const int32_t ab = c->blocks ? c->blocks[3 * i] : 1, tb = c->blocks ? c->blocks[3 * i + 1] : 1;
        if (!valid_rule(ab, c->n_ang[i]) || !valid_rule(tb, c->n_thk[i]))
            return std::string(who) + " " + std::to_string(i) + ": blocks must be >= 1 and increments in [1, 100]";
        if (!(c->R[i] > 0.0) || !(c->a[i] > 0.0) || !(c->b[i] > 0.0) || !std::isfinite(c->J[i]))
            return std::string(who) + " " + std::to_string(i) + ": R, a, b must be positive and J finite";
    }
    return "";
}

}  // namespace

extern "C" ccoil_status ccoil_pairs_zaxis(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                                          const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                                          uint32_t what, uint32_t flags, double *M, double *Fz, ccoil_report *report)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_pairs_zaxis: ctx is NULL");
    if (n_pairs < 0) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_pairs is negative");
    if (n_pairs > INT32_MAX / 2) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_pairs exceeds 2^30 - 1");
    if (what == 0 || (what & ~(uint32_t)(CCOIL_PAIR_M | CCOIL_PAIR_FT)))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "coaxial pairs compute CCOIL_PAIR_M and/or CCOIL_PAIR_FT");
    if ((what & CCOIL_PAIR_M) && !M) return fail(ctx, CCOIL_ERR_ARGUMENT, "M requested but M is NULL");
    if ((what & CCOIL_PAIR_FT) && !Fz) return fail(ctx, CCOIL_ERR_ARGUMENT, "force requested but Fz is NULL");
    std::string msg = check_coils(prim, "primary coil");
    if (msg.empty()) msg = check_coils(sec, "secondary coil");
    if (!msg.empty()) return fail(ctx, CCOIL_ERR_ARGUMENT, msg);
    if (report) std::memset(report, 0, sizeof *report);
    if (n_pairs == 0) return CCOIL_OK;
    if (prim->n == 0 || sec->n == 0) return fail(ctx, CCOIL_ERR_ARGUMENT, "pairs given but a coil array is empty");

    std::vector<ZItemDev> items;
    int64_t evals = 0;
    const int modes[2] = {kZMutual, kZForce};
    for (int m = 0; m < 2; ++m) {
        if (!(what & (m == 0 ? CCOIL_PAIR_M : CCOIL_PAIR_FT))) continue;
        for (int64_t k = 0; k < n_pairs; ++k) {
            const int64_t i = prim_index ? prim_index[k] : k, j = sec_index ? sec_index[k] : k;
            if (i < 0 || i >= prim->n || j < 0 || j >= sec->n)
                return fail(ctx, CCOIL_ERR_ARGUMENT, "pair " + std::to_string(k) + ": coil index out of range");
            ZItemDev it;
            std::memset(&it, 0, sizeof it);
            it.mode = modes[m];
            it.out = (int32_t)(m * n_pairs + k);
            it.a_blk = prim->blocks ? prim->blocks[3 * i] : 1;
            it.t_blk = prim->blocks ? prim->blocks[3 * i + 1] : 1;
            it.r_blk = sec->blocks ? sec->blocks[3 * j + 1] : 1;
            it.a_inc = prim->n_ang[i];
            it.t_inc = prim->n_thk[i];
            it.r_inc = sec->n_thk[j];
            const double t_width = prim->a[i] / it.t_blk, a_width = kPi / it.a_blk, r_width = sec->a[j] / it.r_blk;
            it.t_inner = prim->R[i];
            it.t_width = t_width;
            it.a_width = a_width;
            it.r_inner = sec->R[j];
            it.r_width = r_width;
            it.constant = kMu * prim->J[i] * sec->J[j] * t_width * a_width * r_width * 0.125;
            const double d = sec->pos[3 * j + 2] - prim->pos[3 * i + 2];
            const double hs = sec->b[j] * 0.5, hp = prim->b[i] * 0.5;
            it.z[0] = d + hs + hp;
            it.z[1] = d + hs - hp;
            it.z[2] = d - hs - hp;
            it.z[3] = d - hs + hp;
            for (int q = 0; q < 4; ++q) it.zsq[q] = it.z[q] * it.z[q];
            if (it.mode == kZMutual) {
                if (prim->current[i] == 0.0 || sec->current[j] == 0.0)
                    return fail(ctx, CCOIL_ERR_ARGUMENT, "pair " + std::to_string(k) + ": mutual inductance needs "
                                                         "non-zero currents (M = sum * 2 pi / (I1 I2))");
                it.mul = 2.0 * kPi;
                it.mul2 = 1.0;
                it.div = prim->current[i] * sec->current[j];
            } else {
                it.mul = -2.0 * kPi;
                it.mul2 = 1.0;
                it.div = 1.0;
            }
            it.units = (int64_t)it.r_blk * it.r_inc * it.a_blk * it.a_inc;
            evals += it.units * it.t_blk * it.t_inc;
            items.push_back(it);
        }
    }
    const size_t n_out = (size_t)n_pairs * 2;
    // host results: M then Fz, staged in one device buffer of 2 n_pairs values
    std::vector<double> host(n_out, 0.0);
    const bool dev = flags & CCOIL_DEVICE_POINTERS;
    if (dev) {
        // device outputs are written in place: route M and Fz through one staging buffer, then copy on device
        CC_HIP(ctx, ensure(ctx->zout, sizeof(double) * n_out));
        ccoil_status st = run_items(ctx, items, n_out, flags, nullptr, (double *)ctx->zout.p, evals, report);
        if (st != CCOIL_OK) return st;
        const size_t s1 = sizeof(double) * (size_t)n_pairs;
        if (what & CCOIL_PAIR_M) CC_HIP(ctx, hipMemcpyAsync(M, ctx->zout.p, s1, hipMemcpyDeviceToDevice, ctx->stream));
        if (what & CCOIL_PAIR_FT)
            CC_HIP(ctx, hipMemcpyAsync(Fz, (double *)ctx->zout.p + n_pairs, s1, hipMemcpyDeviceToDevice, ctx->stream));
        CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return CCOIL_OK;
    }
    ccoil_status st = run_items(ctx, items, n_out, flags, host.data(), nullptr, evals, report);
    if (st != CCOIL_OK) return st;
    if (what & CCOIL_PAIR_M) std::memcpy(M, host.data(), sizeof(double) * (size_t)n_pairs);
    if (what & CCOIL_PAIR_FT) std::memcpy(Fz, host.data() + n_pairs, sizeof(double) * (size_t)n_pairs);
    return CCOIL_OK;
}

extern "C" ccoil_status ccoil_self_inductance(ccoil_ctx ctx, const ccoil_coils *coils, const int32_t *radial,
                                              uint32_t flags, double *L, ccoil_report *report)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_self_inductance: ctx is NULL");
    std::string msg = check_coils(coils, "coil");
    if (!msg.empty()) return fail(ctx, CCOIL_ERR_ARGUMENT, msg);
    if (coils->n > INT32_MAX) return fail(ctx, CCOIL_ERR_ARGUMENT, "coil count exceeds 2^31 - 1");
    if (report) std::memset(report, 0, sizeof *report);
    if (coils->n == 0) return CCOIL_OK;
    if (!radial || !L) return fail(ctx, CCOIL_ERR_ARGUMENT, "radial and L must not be NULL");
    std::vector<ZItemDev> items;
    int64_t evals = 0;
    for (int64_t i = 0; i < coils->n; ++i) {
        if (!valid_rule(radial[2 * i], radial[2 * i + 1]))
            return fail(ctx, CCOIL_ERR_ARGUMENT, "coil " + std::to_string(i) + ": radial rule needs blocks >= 1 and "
                                                 "increments in [1, 100]");
        if (coils->current[i] == 0.0)
            return fail(ctx, CCOIL_ERR_ARGUMENT, "coil " + std::to_string(i) + ": self-inductance needs a non-zero "
                                                 "current (L = sum * 2 pi N / (I l))");
        ZItemDev it;
        std::memset(&it, 0, sizeof it);
        it.mode = kZSelf;
        it.out = (int32_t)i;
        it.a_blk = coils->blocks ? coils->blocks[3 * i] : 1;
        it.t_blk = coils->blocks ? coils->blocks[3 * i + 1] : 1;
        it.r_blk = radial[2 * i];
        it.a_inc = coils->n_ang[i];
        it.t_inc = coils->n_thk[i];
        it.r_inc = radial[2 * i + 1];
        const double a = coils->a[i], len = coils->b[i];
        const double t_width = a / it.t_blk, a_width = kPi / it.a_blk, r_unit = 1.0 / it.r_blk;
        it.t_inner = coils->R[i];
        it.t_width = t_width;
        it.a_width = a_width;
        it.r_inner = coils->R[i];
        it.r_width = r_unit * a;   // the radial nodes span the thickness; the constant keeps the unit block
        it.constant = kMu * coils->J[i] * t_width * a_width * r_unit * 2 * 0.125;
        it.z[0] = len;
        it.zsq[0] = len * len;
        it.mul = 2.0 * kPi;
        it.mul2 = coils->turns[i];
        it.div = coils->current[i] * len;
        it.units = (int64_t)it.r_blk * it.r_inc * it.a_blk * it.a_inc;
        evals += it.units * it.t_blk * it.t_inc;
        items.push_back(it);
    }
    return run_items(ctx, items, (size_t)coils->n, flags, L, L, evals, report);
}
