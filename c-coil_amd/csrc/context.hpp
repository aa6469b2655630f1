// Library-private context state shared by the C ABI translation units.
#pragma once

#include "ccoil_hip.h"
#include "kernels.hpp"
#include "pairs.hpp"

#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <string>
#include <vector>

namespace ccoil {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

hipError_t ensure(DevBuf &b, size_t bytes);
void free_buf(DevBuf &b);

// A fixed piece of device memory handed out front to back in 256-byte aligned parts (the no-allocation entry
// point and the plans take their buffers from one instead of the context's grow-only buffers).
struct Arena {
    char *base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    bool fixed = false;   // a fixed-size scratch (the no-allocation entry point): no rule-record pass
    void *take(size_t bytes)
    {
        const size_t off = (used + 255) & ~(size_t)255;
        if (off > cap || bytes > cap - off) return nullptr;
        used = off + bytes;
        return base + off;
    }
    size_t left() const
    {
        const size_t off = (used + 255) & ~(size_t)255;
        return off >= cap ? 0 : cap - off;
    }
};

void rotation(double alpha, double beta, double *T, double *Ti);
void rotation_generators(double alpha, double beta, double *om_a, double *om_b);
std::string pack_coils(const ccoil_coils *c, bool need_len, bool need_freq, std::vector<CoilDev> &out);

}  // namespace ccoil

struct ccoil_ctx_s {
    int device = 0;
    int cu_count = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    // Gauss-Legendre tables (FP64 generated on device, FP32 copies, ring tables)
    double *gl_x = nullptr, *gl_w = nullptr, *gl_cos = nullptr, *gl_omc = nullptr;
    float *gl_xf = nullptr, *gl_wf = nullptr, *gl_cosf = nullptr, *gl_omcf = nullptr;
    double *gl_rcos = nullptr, *gl_rsin = nullptr, *gl_wcos = nullptr;
    float *gl_wcosf = nullptr;
    double *gl_opc = nullptr, *gl_ws = nullptr, *gl_ws2 = nullptr, *gl_wcs = nullptr;
    float *gl_opcf = nullptr, *gl_wcsf = nullptr;
    ccoil::DevBuf coils, pts, outA, outB, outE, outG, fpartial;
    ccoil::DevBuf ncounts, noffsets, nhist, nstate, nkeys, nrecs, nvals, ntemp;   // FP32C near-winding pass
    ccoil::DevBuf fmp, forder, ftemp;                                            // FP32C far field
    int64_t near_cap = 0;                         // records the near buffers hold (grow-only)
    ccoil::NearState *near_host = nullptr;        // pinned copy of the near pass's outcome
    void *scratch = nullptr;                      // the device's no-allocation scratch (noalloc.hip)
    ccoil::DevBuf pprim, psec, ppairs, ppartial, pcounter, pM, pF, pT, pdM;
    ccoil::DevBuf acoils, apts, agB, apartial, agrad;   // field adjoint
    ccoil::DevBuf zitems, zpartial, zout;               // coaxial rules
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // around all compute kernels of a call
    hipEvent_t evm0 = nullptr, evm1 = nullptr; // around the main kernel (field_kernel / pair_kernel)
    hipEvent_t evr0 = nullptr, evr1 = nullptr; // around the record pass's evaluation kernels
    std::map<std::string, std::string> tuning;  // autotuner cache: call key -> chosen launch shape
    std::string tuning_file;                    // where new tuning entries are appended ("" = memory only)
};

namespace ccoil {

// Launch shape of a field call: lanes per point (1, 4 or 16) and coil chunks (grid rows summed in order).
struct FieldShape {
    int split = 1;
    int chunks = 1;
};
// The fixed rule (few points: split rows over lanes; few waves: chunk the coils; large: fill the last round).
// series_only: the call's record pass takes every pair inside the far-field distances, so the launch is the
// series-only kernel (one lane per point).
FieldShape field_shape_rule(ccoil_ctx ctx, int64_t n_points, int64_t n_coils, uint32_t q, int precision,
                            bool series_only = false);

// Everything a field call launches, with its buffers already sized.
struct FieldLaunch {
    FieldParams p;
    int precision = 0;
    int split = 1;
    bool near = false;          // FP32C near pass
    void *scan_temp = nullptr;  // scan storage of the near pass (tile sums)
    size_t scan_bytes = 0;
    void *order_temp = nullptr; // FP32C far field: storage of the visit-order sort
    size_t order_scan_bytes = 0;
    int order_bits = 0;
    size_t far_bytes = 0;       // device bytes of the far-field buffers (moments, order, sort storage)
};
// Far field: the series replaces the rule where a bound on the angular rule's error is below kFarTolerance
// (FP32C) or kFarToleranceFP64 (FP64 with CCOIL_FAR_FIELD) of the coil's field (far_threshold); a spatial visit order
// from kOrderMinPoints points on.
constexpr double kFarTolerance = 1e-11;
constexpr double kFarToleranceFP64 = 1e-14;
constexpr int64_t kOrderMinPoints = 16384;
// Rule records: calls of the series modes with kOrderMinPoints to kRecordMaxPoints points evaluate the rule's pairs
// in the record pass (waves of one coil) instead of beside the series; larger calls are split into pieces of
// kRecordMaxPoints points by ccoil_fields.
constexpr int64_t kRecordMaxPoints = (int64_t)1 << 20;
// Whether a field call of this precision, flags and size has a record pass, and whether it has rule records.
bool field_record_pass(int32_t precision, uint32_t flags, int64_t n_points, bool fixed_arena, bool *rule_records);
double far_threshold(int n_ang, double q, double tol, double ylo);
// The nearest distance, in units of r_max, at which the series may replace a coil's rule on cost, in [kFarMinY, 2].
constexpr double kFarMinY = 1.17;
double far_cost_floor(int n_ang, int n_thk, bool fast, double log_tol);
// Distance, in units of r_max, beyond which the coil's series replaces its rule (0: never): far_threshold at the
// precision's tolerance (fp64: kFarToleranceFP64, else kFarTolerance) from the cost floor on (cost_floor: launches
// with one lane per point or secondary node), else from 2.
double far_distance(const CoilDev &c, bool fp64, bool cost_floor);
// Runs fn(i0, i1) over ranges that cover [0, n) once: on up to 16 host threads for large n, else on the caller.
void host_parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &fn);
// The highest order a series evaluated beyond y r_max can use (multipole_terms there, with a margin; 0 if y is 0).
int far_terms(double y, bool fp64);
// Enqueues the field kernel, the chunk reduction and (FP32C) the whole near pass on s, with no host
// synchronisation and kernel launches and copies only (so it can be captured in a graph); evm0/evm1 (may be NULL)
// bracket the field kernel. The near pass's outcome is copied to host_state (pinned) when given.
void enqueue_fields(hipStream_t s, const FieldLaunch &L, hipEvent_t evm0, hipEvent_t evm1, NearState *host_state,
                    hipEvent_t evr0 = nullptr, hipEvent_t evr1 = nullptr);
// The near pass from the per-coil cursors on (after a grow-on-overflow retry): key scan, fill, eval, add; evr0 / evr1
// (may be NULL) bracket the evaluation kernels.
void enqueue_near_tail(hipStream_t s, const FieldLaunch &L, NearState *host_state, hipEvent_t evr0 = nullptr,
                       hipEvent_t evr1 = nullptr);

ccoil_status fail(ccoil_ctx ctx, ccoil_status code, const std::string &msg);
Tables ctx_tables(ccoil_ctx ctx);
// A call's device buffer of `bytes`: the context's grow-only buffer b (may allocate) when arena is NULL, else the
// arena's next part (never allocates; fails with CCOIL_ERR_ARGUMENT when the arena is too small).
ccoil_status call_buffer(ccoil_ctx ctx, DevBuf &b, Arena *arena, size_t bytes, void **out);

// A field call made ready to launch: packed coils, the launch parameters and their buffers.
struct FieldPrep {
    std::vector<CoilDev> hc;
    FieldLaunch L;
    int64_t n = 0;
    uint32_t q = 0;
    int width = 0;           // doubles per point of the requested outputs
    int64_t nodes = 0;       // quadrature nodes summed over the coils
    bool dev = false;        // points and outputs are device pointers
};
// Validates a field call, uploads its coil records and sizes its buffers: the context's grow-only buffers, or
// parts of arena (then points and outputs must be device pointers). shape: NULL for the launch-shape rule;
// near_cap > 0: the FP32C near pass's record capacity (else every pair when few, or the context's kept capacity).
// A call with no points or no coils returns CCOIL_OK with nothing uploaded.
ccoil_status fields_prepare(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                            uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B, double *E,
                            double *G, Arena *arena, const FieldShape *shape, int64_t near_cap, FieldPrep &out);
// Zero-fills the requested outputs of a call with no coils (device pointers).
void enqueue_zero_outputs(hipStream_t s, const FieldPrep &fp);
// ccoil_fields with its buffers from arena (NULL: the context's buffers) and a launch shape (NULL: the rule). With an
// arena an FP32C near pass that needs more records than the arena holds is an error.
ccoil_status fields_call(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                         uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B, double *E,
                         double *G, ccoil_report *report, Arena *arena, const FieldShape *shape);
ccoil_status near_buffers(ccoil_ctx ctx, FieldLaunch &L, int64_t cap, int width, Arena *arena);
ccoil_status near_outcome(ccoil_ctx ctx, const NearState &st);

// The host side of a pair call: packed coils and the work list, longest work item first.
struct PairBuild {
    std::vector<CoilDev> hp, hs;
    std::vector<PairDev> sorted;
    int64_t items = 0;
    int64_t node_pairs = 0;
    int64_t target = 0;        // primary-node evaluations per lane and work item
    int max_row_groups = 1;
    bool far = false;          // the primaries' far-field series is on (hp[].far_d2 set)
};
// Validates a pair call and builds its work list; target_override > 0 replaces the lane-target rule.
ccoil_status build_pairs(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec, const int64_t *prim_index,
                         const int64_t *sec_index, int64_t n_pairs, const double *ring_offset,
                         const int32_t *half_ring, uint32_t what, int32_t precision, uint32_t flags, double *M,
                         double *F, double *T, double *dM, int64_t target_override, PairBuild &out);
// Uploads a built pair call into bufs (primary coils, secondary coils, work list, partials) and fills the kernel
// parameters (outputs left NULL).
ccoil_status pairs_setup(ccoil_ctx ctx, const PairBuild &pb, uint32_t what, int32_t precision, DevBuf *const bufs[4],
                         PairParams &pp);

// The autotuner: the launch shape of a field call (lanes per point, coil chunks) or the lane target of a pair call,
// from the context's tuning cache, or, the first time a call of that size is seen, by timing the candidates on the
// device (each candidate runs the call itself, so its outputs are written) and caching the fastest. Deterministic
// once cached: the same cache gives the same shape, hence bitwise the same results.
ccoil_status tune_fields(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                         uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B, double *E,
                         double *G, FieldShape &out);
ccoil_status tune_pairs(ccoil_ctx ctx, const PairBuild &rule, uint32_t what, int32_t precision,
                        const std::function<ccoil_status(int64_t, double &)> &run, int64_t &target);
// The tuned lane target of a pair call (rule: its build with the rule's target); tuning runs the call.
ccoil_status pairs_tuned_target(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                                const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                                const double *ring_offset, const int32_t *half_ring, uint32_t what, int32_t precision,
                                uint32_t flags, double *M, double *F, double *T, double *dM, const PairBuild &rule,
                                int64_t &target);
ccoil_status load_tuning_file(ccoil_ctx ctx, const char *path);

// The no-allocation scratch of the current device (noalloc.hip): its device address, and its size.
hipError_t scratch_address(int device, void **out);
size_t scratch_bytes();

#define CC_HIP(ctx, call)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            if (e_ == hipErrorOutOfMemory)                                                                  \
                return ::ccoil::fail(ctx, CCOIL_ERR_OUT_OF_MEMORY, std::string(#call) + ": " + hipGetErrorString(e_)); \
            return ::ccoil::fail(ctx, CCOIL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                                                   \
    } while (0)

}  // namespace ccoil
