// Bridge from the C++ Coil API to the C ABI: per-thread default context, status -> exception mapping,
// and SoA packing of Coil objects into ccoil_coils.
#pragma once

#include "Coil.h"
#include "ccoil_hip.h"

#include <vector>

namespace ccoil {
namespace bridge {

// The calling thread's context on device CCOIL_DEVICE (default 0), created on first use.
ccoil_ctx context();

// Maps a non-OK status to the exception the reference raises in the same situation:
// CCOIL_ERR_ARGUMENT -> std::invalid_argument, CCOIL_ERR_NO_DEVICE -> std::logic_error,
// CCOIL_ERR_OUT_OF_MEMORY -> std::bad_alloc, CCOIL_ERR_DEVICE -> std::runtime_error.
void check(ccoil_status st, ccoil_ctx ctx, const char *what);

// HIP device index of the calling thread's context (CCOIL_DEVICE, default 0).
int device();

// The stream of the calling thread's context (hipStream_t as void*).
void *stream();

int precision();
void setPrecision(int precision);
// The far-field series for this thread's field calls: -1 the precision's default (on in CCOIL_FP32C, off in
// CCOIL_FP64), 0 off, 1 on; fieldFlags() is the matching ccoil_fields flag for the current precision.
int farField();
void setFarField(int mode);
uint32_t fieldFlags();

struct CoilSoA {
    std::vector<int32_t> use_fast, n_ang, n_thk, n_len, blocks;
    std::vector<double> R, a, b, J, current, turns, freq, pos, alpha, beta;

    void add(const Coil &c, const PrecisionArguments &args);
    void add(const Coil &c, const PrecisionArguments &args, const vec3::Vector3 &position, double yAngle,
             double zAngle);
    [[nodiscard]] size_t size() const { return R.size(); }
    [[nodiscard]] ccoil_coils view() const;
};

// Force and torque on point magnetic dipoles in the field of the coils: F = J^T m with J the true Jacobian
// dB_i/dx_j (CCOIL_G_JACOBIAN), T = m x B. One ccoil_fields call for all points.
std::vector<std::pair<vec3::Vector3, vec3::Vector3>> dipoleForceTorque(const CoilSoA &coils,
                                                                       const vec3::Vector3Array &points,
                                                                       const vec3::Vector3Array &moments);

// All requested fields of `coils` at n AoS xyz points in one ccoil_fields call: every non-null output is written in
// place (A, B, E: 3 doubles per point; G: 9, the reference's T*G_local). devicePointers: device buffers on the
// context's device.
void fieldsInto(const CoilSoA &coils, const double *points, size_t n, double *A, double *B, double *E, double *G,
                bool devicePointers);

// Coaxial pairs (Coil::isZAxisCase). THIN/RECTANGULAR pairs take the analytic-length rule of ccoil_pairs_zaxis
// (CalculateMInductanceZAxis.cxx:84-221, CalculateForceZAxis.cxx:82-219); any other pair takes the slow rule:
// |A| (mutual inductance) or B_x (force) of the primary, evaluated by ccoil_fields with the primary arguments, at
// the secondary's r-z grid (CalculateMInductanceZAxis.cxx:17-82, CalculateForceZAxis.cxx:16-80).
double zAxisMutualInductance(const Coil &prim, const Coil &sec, const CoilPairArguments &args);
double zAxisForce(const Coil &prim, const Coil &sec, const CoilPairArguments &args);

// Coil::calculateSelfInductance (CalculateMInductanceZAxis.cxx:224-317) through ccoil_self_inductance.
double selfInductance(const Coil &coil, const CoilPairArguments &args);

}  // namespace bridge
}  // namespace ccoil
