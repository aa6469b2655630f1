// C ABI: plans (include/ccoil_hip.h). A plan is one field or pair call recorded once -- coil records, work list,
// launch shape and every buffer fixed -- and captured into a HIP graph, which each launch replays with a single
// hipGraphLaunch: one submission for the whole call (the pair work queue is one launch of the pair kernel and one
// of its reduction), no host work and no host synchronisation per launch.
#include "context.hpp"

#include <algorithm>
#include <cstring>

using namespace ccoil;

struct ccoil_plan_s {
    ccoil_ctx ctx = nullptr;
    bool empty = false;                 // nothing to evaluate (no points or no pairs): launches do nothing
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool launched = false;
    DevBuf mem;                         // field plans: every buffer of the call, handed out by an Arena
    DevBuf pbufs[4];                    // pair plans: primary coils, secondary coils, work list, partials
    NearState *host_state = nullptr;    // field plans with a record pass: the last launch's outcome (pinned)
    bool near = false;
    bool fp32c = false;
    int64_t near_cap = 0;
    int64_t node_evaluations = 0;
    int32_t split = 1;
    int32_t block = 0;
};

namespace {

ccoil_status plan_fail(ccoil_plan plan, ccoil_status st)
{
    ccoil_plan_destroy(plan);
    return st;
}

// Captures what enqueue puts on a stream into the plan's graph and instantiates it.
template <typename F> ccoil_status capture(ccoil_plan plan, F &&enqueue)
{
    ccoil_ctx ctx = plan->ctx;
    hipStream_t cs = nullptr;
    CC_HIP(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        enqueue(cs);
        const hipError_t le = hipGetLastError();
        e = hipStreamEndCapture(cs, &plan->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    CC_HIP(ctx, e);
    return CCOIL_OK;
}

ccoil_status new_plan(ccoil_ctx ctx, ccoil_plan *out, ccoil_plan &plan)
{
    *out = nullptr;
    plan = new ccoil_plan_s();
    plan->ctx = ctx;
    CC_HIP(ctx, hipSetDevice(ctx->device));
    CC_HIP(ctx, hipEventCreate(&plan->ev0));
    CC_HIP(ctx, hipEventCreate(&plan->ev1));
    return CCOIL_OK;
}

}  // namespace

extern "C" {

ccoil_status ccoil_fields_plan(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                               uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B,
                               double *E, double *G, ccoil_plan *out)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_fields_plan: ctx is NULL");
    if (!out) return fail(ctx, CCOIL_ERR_ARGUMENT, "ccoil_fields_plan: out is NULL");
    *out = nullptr;
    if (!(flags & CCOIL_DEVICE_POINTERS))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "a plan takes device pointers (CCOIL_DEVICE_POINTERS)");
    // the launch shape: tuned (the tuning runs the call) or the rule
    FieldShape shape;
    const FieldShape *sh = nullptr;
    if (flags & CCOIL_AUTOTUNE) {
        const ccoil_status st = tune_fields(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G,
                                            shape);
        if (st != CCOIL_OK) return st;
        sh = &shape;
    }
    // a call with a record pass (FP32C's near pairs, the series modes' rule records): one eager call counts the
    // records of these points; the plan holds a quarter more
    int64_t cap = 0;
    if (field_record_pass(precision, flags, n_points, false, nullptr) && n_points > 0) {
        ccoil_report rep;
        const ccoil_status st = fields_call(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G,
                                            &rep, nullptr, sh);
        if (st != CCOIL_OK) return st;
        const int64_t total = (int64_t)ctx->near_host->total;
        cap = total + total / 4 + 4096;
    }
    ccoil_plan plan;
    ccoil_status st = new_plan(ctx, out, plan);
    if (st != CCOIL_OK) return plan_fail(plan, st);
    // validation and sizes with the context's buffers, then every buffer once more out of the plan's own memory
    FieldPrep probe;
    st = fields_prepare(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G, nullptr, sh, cap,
                        probe);
    if (st != CCOIL_OK) return plan_fail(plan, st);
    plan->node_evaluations = probe.nodes * probe.n;
    if (probe.n == 0) {
        plan->empty = true;
        *out = plan;
        return CCOIL_OK;
    }
    const FieldParams &pp = probe.L.p;
    const size_t nc = probe.hc.size();
    size_t bytes = (sizeof(CoilDev) + sizeof(uint32_t)) * nc + 4096;
    if (pp.n_chunks > 1) bytes += sizeof(double) * (size_t)pp.partial_width * (size_t)probe.n * (size_t)pp.n_chunks + 256;
    bytes += probe.L.far_bytes;
    if (probe.L.near) {
        bytes += 2 * sizeof(int64_t) * (size_t)probe.n + 2 * sizeof(uint32_t) * ((size_t)pp.n_keys + 2) + sizeof(NearState) +
                 probe.L.scan_bytes + (size_t)pp.near_cap * (sizeof(uint32_t) + sizeof(uint64_t) +
                                                             sizeof(double) * (size_t)probe.width) + 8 * 256;
        plan->near = true;
        plan->fp32c = precision == CCOIL_FP32C;
        plan->near_cap = pp.near_cap;
        if (hipHostMalloc((void **)&plan->host_state, sizeof(NearState), hipHostMallocDefault) != hipSuccess) {
            plan->host_state = nullptr;
            fail(ctx, CCOIL_ERR_OUT_OF_MEMORY, "plan: hipHostMalloc of the near-pass state failed");
            return plan_fail(plan, CCOIL_ERR_OUT_OF_MEMORY);
        }
        std::memset(plan->host_state, 0, sizeof(NearState));
    }
    {
        const hipError_t e = ensure(plan->mem, bytes);
        if (e != hipSuccess) {
            fail(ctx, CCOIL_ERR_OUT_OF_MEMORY, std::string("plan buffers: ") + hipGetErrorString(e));
            return plan_fail(plan, CCOIL_ERR_OUT_OF_MEMORY);
        }
    }
    Arena arena;
    arena.base = (char *)plan->mem.p;
    arena.cap = plan->mem.cap;
    FieldPrep fp;   // the same shape (tuned, or the rule on the same inputs) and capacity as the probe
    st = fields_prepare(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G, &arena, sh,
                        probe.L.near ? pp.near_cap : 0, fp);
    if (st != CCOIL_OK) return plan_fail(plan, st);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)   // the coil records are in place before any replay
        return plan_fail(plan, fail(ctx, CCOIL_ERR_DEVICE, "plan: uploading the coil records failed"));
    plan->split = fp.L.split;
    plan->block = kFieldBlock;
    if (fp.hc.empty()) st = capture(plan, [&](hipStream_t s) { enqueue_zero_outputs(s, fp); });
    else st = capture(plan, [&](hipStream_t s) { enqueue_fields(s, fp.L, nullptr, nullptr, plan->host_state); });
    if (st != CCOIL_OK) return plan_fail(plan, st);
    *out = plan;
    return CCOIL_OK;
}

ccoil_status ccoil_pairs_plan(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                              const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                              const double *ring_offset, const int32_t *half_ring, uint32_t what, int32_t precision,
                              uint32_t flags, double *M, double *F, double *T, double *dM, ccoil_plan *out)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_pairs_plan: ctx is NULL");
    if (!out) return fail(ctx, CCOIL_ERR_ARGUMENT, "ccoil_pairs_plan: out is NULL");
    *out = nullptr;
    if (!(flags & CCOIL_DEVICE_POINTERS))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "a plan takes device pointers (CCOIL_DEVICE_POINTERS)");
    PairBuild pb;
    ccoil_status st = build_pairs(ctx, prim, sec, prim_index, sec_index, n_pairs, ring_offset, half_ring, what,
                                  precision, flags, M, F, T, dM, 0, pb);
    if (st != CCOIL_OK) return st;
    if (n_pairs > 0 && (flags & CCOIL_AUTOTUNE)) {
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
    ccoil_plan plan;
    st = new_plan(ctx, out, plan);
    if (st != CCOIL_OK) return plan_fail(plan, st);
    plan->node_evaluations = pb.node_pairs;
    if (n_pairs == 0) {
        plan->empty = true;
        *out = plan;
        return CCOIL_OK;
    }
    PairParams pp;
    DevBuf *const bufs[4] = {&plan->pbufs[0], &plan->pbufs[1], &plan->pbufs[2], &plan->pbufs[3]};
    st = pairs_setup(ctx, pb, what, precision, bufs, pp);
    if (st != CCOIL_OK) return plan_fail(plan, st);
    pp.M = M;
    pp.F = F;
    pp.T = T;
    pp.dM = dM;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)   // the coil records and the work list are in place
        return plan_fail(plan, fail(ctx, CCOIL_ERR_DEVICE, "plan: uploading the pair work list failed"));
    plan->split = pb.max_row_groups;
    plan->block = 64;
    st = capture(plan, [&](hipStream_t s) {
        launch_pairs(s, pp);
        launch_pair_reduce(s, pp);
    });
    if (st != CCOIL_OK) return plan_fail(plan, st);
    *out = plan;
    return CCOIL_OK;
}

ccoil_status ccoil_plan_launch(ccoil_plan plan, void *stream)
{
    if (!plan) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_plan_launch: plan is NULL");
    ccoil_ctx ctx = plan->ctx;
    if (plan->empty) return CCOIL_OK;
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    CC_HIP(ctx, hipSetDevice(ctx->device));
    CC_HIP(ctx, hipEventRecord(plan->ev0, s));
    CC_HIP(ctx, hipGraphLaunch(plan->exec, s));
    CC_HIP(ctx, hipEventRecord(plan->ev1, s));
    plan->launched = true;
    return CCOIL_OK;
}

ccoil_status ccoil_plan_synchronize(ccoil_plan plan, ccoil_report *report)
{
    if (!plan) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_plan_synchronize: plan is NULL");
    ccoil_ctx ctx = plan->ctx;
    if (report) std::memset(report, 0, sizeof *report);
    if (plan->empty || !plan->launched) return CCOIL_OK;
    CC_HIP(ctx, hipSetDevice(ctx->device));
    CC_HIP(ctx, hipEventSynchronize(plan->ev1));
    if (plan->near) {
        const ccoil_status st = near_outcome(ctx, *plan->host_state);
        if (st != CCOIL_OK) return st;
        if (plan->host_state->overflow)
            return fail(ctx, CCOIL_ERR_CAPACITY, "field plan: this launch found " +
                        std::to_string(plan->host_state->total) + " record-pass (coil, point) pairs (near-winding "
                        "or rule records), the plan holds " + std::to_string(plan->near_cap) +
                        "; its outputs lack those contributions; make a new plan");
    }
    if (report) {
        float ms = 0.f;
        CC_HIP(ctx, hipEventElapsedTime(&ms, plan->ev0, plan->ev1));
        report->node_evaluations = plan->node_evaluations;
        report->near_pairs = plan->near && plan->fp32c ? (int64_t)plan->host_state->fp64 : 0;
        report->kernel_ms = ms;
        report->main_kernel_ms = ms;
        report->block_size = plan->block;
        report->split = plan->split;
    }
    return CCOIL_OK;
}

ccoil_status ccoil_plan_destroy(ccoil_plan plan)
{
    if (!plan) return CCOIL_OK;
    (void)hipSetDevice(plan->ctx->device);
    if (plan->launched) (void)hipEventSynchronize(plan->ev1);
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->ev0) (void)hipEventDestroy(plan->ev0);
    if (plan->ev1) (void)hipEventDestroy(plan->ev1);
    free_buf(plan->mem);
    for (auto &b : plan->pbufs) free_buf(b);
    if (plan->host_state) (void)hipHostFree(plan->host_state);
    delete plan;
    return CCOIL_OK;
}

}  // extern "C"
