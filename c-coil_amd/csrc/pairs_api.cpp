// C ABI: ccoil_pairs (mutual inductance, force and torque of many coil pairs, include/ccoil_hip.h).
#include "context.hpp"
#include "multipole.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

using namespace ccoil;

namespace {

// Primary-node evaluations per lane and work item. Large batches give every lane about 8192 (each item's
// fixed costs, the wave reduction and the partial, stay small); a batch with too little work to give the 1,024
// SIMDs about 4 waves each at that size lowers it, down to 256, so that a single pair or a few pairs still spread
// over the chip instead of running one long chain per lane. Powers of two only.
int64_t lane_target(int64_t node_pairs)
{
    int64_t t = 8192;
    while (t > 256 && node_pairs < (int64_t)64 * 4096 * t) t >>= 1;
    return t;
}

// Work decomposition of one pair for a lane target: secondary nodes per item (64 lanes times nodes per lane)
// and the number of groups the primary's thickness rows are split into. Small primaries take several secondary
// nodes per lane; a primary with more nodes than the target splits its rows.
void decompose(int64_t target, int n_ang, int n_thk, int &nodes_per_item, int &row_groups)
{
    const int64_t prim_nodes = (int64_t)n_ang * n_thk;
    if (prim_nodes <= target) {
        const int64_t per_lane = std::max<int64_t>(1, std::min<int64_t>(64, (target + prim_nodes - 1) / prim_nodes));
        nodes_per_item = (int)(64 * per_lane);
        row_groups = 1;
    } else {
        nodes_per_item = 64;
        const int64_t rows_per_group = std::max<int64_t>(1, target / n_ang);
        row_groups = (int)std::min<int64_t>(n_thk, (n_thk + rows_per_group - 1) / rows_per_group);
    }
}

// The cost of the primary's series at one secondary node, in primary-rule nodes (about 30 terms of 13 flops
// against 20-40 flops per rule node), for the work decomposition of pairs that take the series at every node.
constexpr int kSeriesNodes = 16;

// Every secondary node lies beyond the primary's far-field distance: the secondary's nodes lie within
// |(R + a, b / 2)| of its centre.
bool series_only(const CoilDev &p, const CoilDev &s)
{
    if (!(p.far_d2 > 0.0)) return false;
    const double d[3] = {s.pos[0] - p.pos[0], s.pos[1] - p.pos[1], s.pos[2] - p.pos[2]};
    const double gap = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - std::hypot(s.R + s.a, 0.5 * s.b);
    return gap > 0.0 && gap * gap > p.far_d2;
}

}  // namespace

namespace ccoil {

ccoil_status build_pairs(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec, const int64_t *prim_index,
                         const int64_t *sec_index, int64_t n_pairs, const double *ring_offset,
                         const int32_t *half_ring, uint32_t what, int32_t precision, uint32_t flags, double *M,
                         double *F, double *T, double *dM, int64_t target_override, PairBuild &out)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_pairs: ctx is NULL");
    if (n_pairs < 0) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_pairs is negative");
    if (n_pairs > INT32_MAX) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_pairs exceeds 2^31 - 1");
    if (what == 0 || (what & ~(uint32_t)(CCOIL_PAIR_M | CCOIL_PAIR_FT | CCOIL_PAIR_DM)))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "bad pair quantity mask");
    if (precision != CCOIL_FP64 && precision != CCOIL_FP32C && precision != CCOIL_FP32 &&
        precision != CCOIL_FP64_REFERENCE && precision != CCOIL_FP64_IEEE)
        return fail(ctx, CCOIL_ERR_ARGUMENT, "unknown precision");
    if ((what & CCOIL_PAIR_M) && !M) return fail(ctx, CCOIL_ERR_ARGUMENT, "M requested but M is NULL");
    if ((what & CCOIL_PAIR_FT) && (!F || !T)) return fail(ctx, CCOIL_ERR_ARGUMENT, "F/T requested but F or T is NULL");
    if ((what & CCOIL_PAIR_DM) && !dM) return fail(ctx, CCOIL_ERR_ARGUMENT, "dM requested but dM is NULL");
    if ((what & CCOIL_PAIR_DM) && (precision == CCOIL_FP32 || precision == CCOIL_FP64_REFERENCE))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "CCOIL_PAIR_DM needs CCOIL_FP64, CCOIL_FP64_IEEE or CCOIL_FP32C");
    std::vector<CoilDev> &hp = out.hp, &hs = out.hs;
    out.sorted.clear();
    out.items = out.node_pairs = 0;
    out.max_row_groups = 1;
    std::string msg = pack_coils(prim, false, false, hp);
    if (msg.empty()) msg = pack_coils(sec, true, false, hs);
    if (!msg.empty()) return fail(ctx, CCOIL_ERR_ARGUMENT, msg);
    if (!prim->current || !prim->turns) return fail(ctx, CCOIL_ERR_ARGUMENT, "primary coils need current and turns");
    if (n_pairs == 0) return CCOIL_OK;
    if (hp.empty() || hs.empty()) return fail(ctx, CCOIL_ERR_ARGUMENT, "pairs given but a coil array is empty");
    // the primaries' far-field series: FP32C by default, FP64 with CCOIL_FAR_FIELD; never for the gradient of M
    out.far = !(what & kPairGrad) && ((precision == CCOIL_FP32C && !(flags & CCOIL_DIRECT)) ||
                                      (precision == CCOIL_FP64 && (flags & CCOIL_FAR_FIELD)));
    if (out.far)
        host_parallel_for((int64_t)hp.size(), [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i) {
                CoilDev &c = hp[(size_t)i];
                const double y = far_distance(c, precision == CCOIL_FP64, true);
                c.far_d2 = y > 0.0 ? (y * c.rmax) * (y * c.rmax) : 0.0;
                c.mp_terms = far_terms(y, precision == CCOIL_FP64);
            }
        });

    // Pair table, longest work item first (LPT order of the persistent queue).
    std::vector<PairDev> pd((size_t)n_pairs);
    int64_t node_pairs = 0;
    const bool ref_scale = flags & CCOIL_FORCE_REFERENCE_SCALE;
    for (int64_t k = 0; k < n_pairs; ++k) {
        const int64_t pi = prim_index ? prim_index[k] : k, si = sec_index ? sec_index[k] : k;
        if (pi < 0 || pi >= (int64_t)hp.size() || si < 0 || si >= (int64_t)hs.size())
            return fail(ctx, CCOIL_ERR_ARGUMENT, "pair " + std::to_string(k) + ": coil index out of range");
        const CoilDev &p = hp[(size_t)pi], &s = hs[(size_t)si];
        node_pairs += (int64_t)s.blk_ang * s.n_ang * s.blk_thk * s.n_thk * s.blk_len * s.n_len * p.n_ang * p.n_thk;
    }
    const int64_t target = target_override > 0 ? target_override : lane_target(node_pairs);
    out.target = target;
    int max_row_groups = 1;
    for (int64_t k = 0; k < n_pairs; ++k) {
        const int64_t pi = prim_index ? prim_index[k] : k, si = sec_index ? sec_index[k] : k;
        const CoilDev &p = hp[(size_t)pi], &s = hs[(size_t)si];
        PairDev d;
        std::memset(&d, 0, sizeof d);
        d.prim = (int32_t)pi;
        d.sec = (int32_t)si;
        d.out = (int32_t)k;
        d.n_nodes = (int64_t)s.blk_ang * s.n_ang * s.blk_thk * s.n_thk * s.blk_len * s.n_len;
        if (series_only(p, s)) decompose(target, kSeriesNodes, 1, d.nodes_per_item, d.row_groups);
        else decompose(target, p.n_ang, p.n_thk, d.nodes_per_item, d.row_groups);
        max_row_groups = std::max(max_row_groups, d.row_groups);
        d.ring_offset = ring_offset ? ring_offset[k] : 0.0;
        d.half_ring = half_ring ? (half_ring[k] ? 1 : 0) : 0;
        const double blocks = (double)s.blk_ang * s.blk_thk * s.blk_len;
        if (p.current == 0.0 && (what & (CCOIL_PAIR_M | CCOIL_PAIR_DM)))
            return fail(ctx, CCOIL_ERR_ARGUMENT, "pair " + std::to_string(k) + ": primary current is zero");
        d.m_scale = 2.0 * kPi * s.turns / (p.current * blocks);
        d.f_scale = ref_scale ? 2.0 * kPi * s.turns / (p.current * blocks) : 2.0 * kPi * s.turns * s.current / blocks;
        pd[(size_t)k] = d;
    }
    // cost of one work item: its secondary nodes times the primary nodes of its row group
    std::vector<int64_t> cost((size_t)n_pairs);
    for (int64_t k = 0; k < n_pairs; ++k) {
        const PairDev &d = pd[(size_t)k];
        const CoilDev &p = hp[(size_t)d.prim];
        const int64_t rows = (p.n_thk + d.row_groups - 1) / d.row_groups;
        cost[(size_t)k] = series_only(p, hs[(size_t)d.sec])
                              ? std::min<int64_t>(d.nodes_per_item, d.n_nodes) * kSeriesNodes
                              : std::min<int64_t>(d.nodes_per_item, d.n_nodes) * p.n_ang * rows * (p.use_fast ? 2 : 1);
    }
    std::vector<int32_t> order((size_t)n_pairs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cost[(size_t)x] > cost[(size_t)y]; });
    std::vector<PairDev> &sorted = out.sorted;
    sorted.resize((size_t)n_pairs);
    // items are counted below; one workgroup per item must fit a 1-D grid
    int64_t items = 0;
    for (int64_t k = 0; k < n_pairs; ++k) {
        PairDev d = pd[(size_t)order[(size_t)k]];
        d.first_item = items;
        items += (d.n_nodes + d.nodes_per_item - 1) / d.nodes_per_item * d.row_groups;
        sorted[(size_t)k] = d;
    }

    if (items > INT32_MAX)
        return fail(ctx, CCOIL_ERR_ARGUMENT, "pair batch needs more than 2^31 - 1 work items; split the batch");
    out.items = items;
    out.node_pairs = node_pairs;
    out.max_row_groups = max_row_groups;
    return CCOIL_OK;
}

}  // namespace ccoil

namespace ccoil {

ccoil_status pairs_setup(ccoil_ctx ctx, const PairBuild &pb, uint32_t what, int32_t precision, DevBuf *const bufs[4],
                         PairParams &pp)
{
    CC_HIP(ctx, hipSetDevice(ctx->device));
    const std::vector<CoilDev> &hp = pb.hp, &hs = pb.hs;
    const std::vector<PairDev> &sorted = pb.sorted;
    // primary records, then (series) the moments: kMpHeader + kMpCoil per primary
    const size_t prim_bytes = (sizeof(CoilDev) * hp.size() + 255) & ~(size_t)255;
    const size_t mp_bytes = pb.far ? sizeof(double) * ((size_t)kMpHeader + (size_t)kMpCoil * hp.size()) : 0;
    CC_HIP(ctx, ensure(*bufs[0], prim_bytes + mp_bytes));
    CC_HIP(ctx, ensure(*bufs[1], sizeof(CoilDev) * hs.size()));
    CC_HIP(ctx, ensure(*bufs[2], sizeof(PairDev) * sorted.size()));
    const int terms = (what & kPairGrad) ? kPairGradTerms : kPairTerms;
    CC_HIP(ctx, ensure(*bufs[3], sizeof(double) * terms * (size_t)pb.items));
    CC_HIP(ctx, hipMemcpyAsync(bufs[0]->p, hp.data(), sizeof(CoilDev) * hp.size(), hipMemcpyHostToDevice, ctx->stream));
    CC_HIP(ctx, hipMemcpyAsync(bufs[1]->p, hs.data(), sizeof(CoilDev) * hs.size(), hipMemcpyHostToDevice, ctx->stream));
    CC_HIP(ctx, hipMemcpyAsync(bufs[2]->p, sorted.data(), sizeof(PairDev) * sorted.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    std::memset(&pp, 0, sizeof pp);
    pp.prim = (const CoilDev *)bufs[0]->p;
    pp.mp = pb.far && !(what & kPairGrad) ? (double *)((char *)bufs[0]->p + prim_bytes) : nullptr;
    pp.n_prim = (int32_t)hp.size();
    pp.sec = (const CoilDev *)bufs[1]->p;
    pp.pairs = (const PairDev *)bufs[2]->p;
    pp.n_pairs = (int32_t)sorted.size();
    pp.n_items = pb.items;
    pp.partial = (double *)bufs[3]->p;
    pp.what = what;
    pp.precision = precision;
    pp.tab = ctx_tables(ctx);
    return CCOIL_OK;
}

// One pair call from a built work list, on the context's buffers; synchronous.
static ccoil_status pairs_run(ccoil_ctx ctx, const PairBuild &pb, uint32_t what, int32_t precision, uint32_t flags,
                              double *M, double *F, double *T, double *dM, ccoil_report *report)
{
    const int64_t n_pairs = (int64_t)pb.sorted.size();
    PairParams pp;
    DevBuf *const bufs[4] = {&ctx->pprim, &ctx->psec, &ctx->ppairs, &ctx->ppartial};
    ccoil_status st = pairs_setup(ctx, pb, what, precision, bufs, pp);
    if (st != CCOIL_OK) return st;
    const bool dev = flags & CCOIL_DEVICE_POINTERS;
    const size_t s1 = sizeof(double) * (size_t)n_pairs, s3 = 3 * s1;
    if (dev) {
        pp.M = M;
        pp.F = F;
        pp.T = T;
        pp.dM = dM;
    } else {
        if (what & CCOIL_PAIR_DM) { CC_HIP(ctx, ensure(ctx->pdM, s3)); pp.dM = (double *)ctx->pdM.p; }
        if (what & CCOIL_PAIR_M) { CC_HIP(ctx, ensure(ctx->pM, s1)); pp.M = (double *)ctx->pM.p; }
        if (what & CCOIL_PAIR_FT) {
            CC_HIP(ctx, ensure(ctx->pF, s3));
            CC_HIP(ctx, ensure(ctx->pT, s3));
            pp.F = (double *)ctx->pF.p;
            pp.T = (double *)ctx->pT.p;
        }
    }
    CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    launch_pairs(ctx->stream, pp);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->evm1, ctx->stream));
    launch_pair_reduce(ctx->stream, pp);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!dev) {
        if (what & CCOIL_PAIR_M) CC_HIP(ctx, hipMemcpyAsync(M, pp.M, s1, hipMemcpyDeviceToHost, ctx->stream));
        if (what & CCOIL_PAIR_FT) {
            CC_HIP(ctx, hipMemcpyAsync(F, pp.F, s3, hipMemcpyDeviceToHost, ctx->stream));
            CC_HIP(ctx, hipMemcpyAsync(T, pp.T, s3, hipMemcpyDeviceToHost, ctx->stream));
        }
        if (what & CCOIL_PAIR_DM) CC_HIP(ctx, hipMemcpyAsync(dM, pp.dM, s3, hipMemcpyDeviceToHost, ctx->stream));
    }
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (report) {
        float ms = 0.f, mms = 0.f;
        CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        CC_HIP(ctx, hipEventElapsedTime(&mms, ctx->ev0, ctx->evm1));
        report->node_evaluations = pb.node_pairs;
        report->kernel_ms = ms;
        report->main_kernel_ms = mms;
        report->block_size = 64;
        report->split = pb.max_row_groups;
    }
    return CCOIL_OK;
}

ccoil_status pairs_tuned_target(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                                const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                                const double *ring_offset, const int32_t *half_ring, uint32_t what, int32_t precision,
                                uint32_t flags, double *M, double *F, double *T, double *dM, const PairBuild &rule,
                                int64_t &target)
{
    auto run = [&](int64_t t, double &ms) -> ccoil_status {
        PairBuild pb;
        ccoil_status st = build_pairs(ctx, prim, sec, prim_index, sec_index, n_pairs, ring_offset, half_ring, what,
                                      precision, flags, M, F, T, dM, t, pb);
        if (st != CCOIL_OK) return st;
        ccoil_report rep;
        st = pairs_run(ctx, pb, what, precision, flags, M, F, T, dM, &rep);
        ms = rep.main_kernel_ms;
        return st;
    };
    return tune_pairs(ctx, rule, what, precision, run, target);
}

}  // namespace ccoil

extern "C" ccoil_status ccoil_pairs(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                                    const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                                    const double *ring_offset, const int32_t *half_ring, uint32_t what,
                                    int32_t precision, uint32_t flags, double *M, double *F, double *T,
                                    double *dM, ccoil_report *report)
{
    if (report) std::memset(report, 0, sizeof *report);
    PairBuild pb;
    ccoil_status st = build_pairs(ctx, prim, sec, prim_index, sec_index, n_pairs, ring_offset, half_ring, what,
                                  precision, flags, M, F, T, dM, 0, pb);
    if (st != CCOIL_OK || n_pairs == 0) return st;
    if (flags & CCOIL_AUTOTUNE) {
        int64_t target = pb.target;
        st = pairs_tuned_target(ctx, prim, sec, prim_index, sec_index, n_pairs, ring_offset, half_ring, what,
                                precision, flags, M, F, T, dM, pb, target);
        if (st != CCOIL_OK) return st;
        if (target != pb.target) {
            st = build_pairs(ctx, prim, sec, prim_index, sec_index, n_pairs, ring_offset, half_ring, what, precision,
                             flags, M, F, T, dM, target, pb);
            if (st != CCOIL_OK) return st;
        }
    }
    return pairs_run(ctx, pb, what, precision, flags, M, F, T, dM, report);
}

extern "C" ccoil_status ccoil_pairs_grad(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                                         const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                                         const double *ring_offset, const int32_t *half_ring, int32_t precision,
                                         uint32_t flags, double *M, double *grad, ccoil_report *report)
{
    if (report) std::memset(report, 0, sizeof *report);
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_pairs_grad: ctx is NULL");
    if (precision != CCOIL_FP64 && precision != CCOIL_FP64_IEEE && precision != CCOIL_FP32C)
        return fail(ctx, CCOIL_ERR_ARGUMENT, "ccoil_pairs_grad supports CCOIL_FP64, CCOIL_FP64_IEEE and CCOIL_FP32C");
    if (!grad) return fail(ctx, CCOIL_ERR_ARGUMENT, "grad is NULL");
    PairBuild pb;
    // the rule at every node: the series carries no derivatives with respect to the primary's R, a and b
    const uint32_t rule_flags = (flags & ~(uint32_t)CCOIL_FAR_FIELD) | (uint32_t)CCOIL_DIRECT;
    ccoil_status st = build_pairs(ctx, prim, sec, prim_index, sec_index, n_pairs, ring_offset, half_ring,
                                  CCOIL_PAIR_M, precision, rule_flags, M ? M : grad, nullptr, nullptr, nullptr, 0, pb);
    if (st != CCOIL_OK || n_pairs == 0) return st;
    PairParams pp;
    DevBuf *const bufs[4] = {&ctx->pprim, &ctx->psec, &ctx->ppairs, &ctx->ppartial};
    const uint32_t what = CCOIL_PAIR_M | kPairGrad;
    st = pairs_setup(ctx, pb, what, precision, bufs, pp);
    if (st != CCOIL_OK) return st;
    const bool dev = flags & CCOIL_DEVICE_POINTERS;
    const size_t s1 = sizeof(double) * (size_t)n_pairs, sg = kPairGradOut * s1;
    if (dev) {
        pp.M = M;
        pp.grad = grad;
    } else {
        CC_HIP(ctx, ensure(ctx->pdM, sg));
        pp.grad = (double *)ctx->pdM.p;
        if (M) {
            CC_HIP(ctx, ensure(ctx->pM, s1));
            pp.M = (double *)ctx->pM.p;
        }
    }
    CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    launch_pairs(ctx->stream, pp);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->evm1, ctx->stream));
    launch_pair_reduce(ctx->stream, pp);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!dev) {
        CC_HIP(ctx, hipMemcpyAsync(grad, pp.grad, sg, hipMemcpyDeviceToHost, ctx->stream));
        if (M) CC_HIP(ctx, hipMemcpyAsync(M, pp.M, s1, hipMemcpyDeviceToHost, ctx->stream));
    }
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (report) {
        float ms = 0.f, mms = 0.f;
        CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        CC_HIP(ctx, hipEventElapsedTime(&mms, ctx->ev0, ctx->evm1));
        report->node_evaluations = pb.node_pairs;
        report->kernel_ms = ms;
        report->main_kernel_ms = mms;
        report->block_size = 64;
        report->split = pb.max_row_groups;
    }
    return CCOIL_OK;
}
