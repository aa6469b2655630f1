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
    ccoil::DevBuf ncounts, noffsets, nkeys, nkeys2, nrecs, nrecs2, nvals, ntemp;   // FP32C near-winding pass
    ccoil::DevBuf pprim, psec, ppairs, ppartial, pcounter, pM, pF, pT, pdM;
    ccoil::DevBuf acoils, apts, agB, apartial, agrad;   // field adjoint
    ccoil::DevBuf zitems, zpartial, zout;               // coaxial rules
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // around all compute kernels of a call
    hipEvent_t evm0 = nullptr, evm1 = nullptr; // around the main kernel (field_kernel / pair_kernel)
};

namespace ccoil {

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
