// Library-private context state shared by the C ABI translation units.
#pragma once

#include "ccoil_hip.h"
#include "kernels.hpp"
#include "pairs.hpp"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace ccoil {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

hipError_t ensure(DevBuf &b, size_t bytes);
void free_buf(DevBuf &b);
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
    int64_t near_cap = 0;                         // records the near buffers hold (grow-only)
    ccoil::NearState *near_host = nullptr;        // pinned copy of the near pass's outcome
    ccoil::DevBuf pprim, psec, ppairs, ppartial, pcounter, pM, pF, pT, pdM;
    ccoil::DevBuf acoils, apts, agB, apartial, agrad;   // field adjoint
    ccoil::DevBuf zitems, zpartial, zout;               // coaxial rules
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // around all compute kernels of a call
    hipEvent_t evm0 = nullptr, evm1 = nullptr; // around the main kernel (field_kernel / pair_kernel)
};

namespace ccoil {

// Launch shape of a field call: lanes per point (1, 4 or 16) and coil chunks (grid rows summed in order).
struct FieldShape {
    int split = 1;
    int chunks = 1;
};
// The fixed rule (few points: split rows over lanes; few waves: chunk the coils; large: fill the last round).
FieldShape field_shape_rule(ccoil_ctx ctx, int64_t n_points, int64_t n_coils, uint32_t q, int precision);

// Everything a field call launches, with its buffers already sized.
struct FieldLaunch {
    FieldParams p;
    int precision = 0;
    int split = 1;
    bool near = false;          // FP32C near pass
    void *scan_temp = nullptr;  // hipCUB scan storage of the near pass
    size_t scan_bytes = 0;
};
// Enqueues the field kernel, the chunk reduction and (FP32C) the whole near pass on s, with no host
// synchronisation; evm0/evm1 (may be NULL) bracket the field kernel. The near pass's outcome is copied to
// host_state (pinned) when given.
void enqueue_fields(hipStream_t s, const FieldLaunch &L, hipEvent_t evm0, hipEvent_t evm1, NearState *host_state);
// The near pass from the per-coil cursors on (after a grow-on-overflow retry): key scan, fill, eval, add.
void enqueue_near_tail(hipStream_t s, const FieldLaunch &L, NearState *host_state);

ccoil_status fail(ccoil_ctx ctx, ccoil_status code, const std::string &msg);
Tables ctx_tables(ccoil_ctx ctx);

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
