/*
 * ccoil_oracle.h -- FP64 CPU restatement of C-Coil's field, planning and pair algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. This library is what the tests check the HIP product path against: only the test
 * suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and then only to
 * compare against or to time as the CPU baseline. The product library (libccoil_hip.so and the
 * c_coil module) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference root).
 * Parity status of each part is stated next to it ("pinned by ..." or "parity unpinned").
 */
#ifndef CCOIL_ORACLE_H
#define CCOIL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_RECTANGULAR = 0, ORC_THIN = 1, ORC_FLAT = 2, ORC_FILAMENT = 3 };

/* A coil after construction, src/Coil/Coil.cxx:22-44 + src/Coil/Utils/CalculateAttributes.cxx:50-78. */
typedef struct {
    double R, a, b;          /* inner radius, thickness (clamped), length (clamped) */
    int32_t N;               /* turns */
    double I;                /* current */
    double J;                /* current density N*I/(a*b) */
    double freq;             /* sine frequency (0 = DC) */
    int32_t type;            /* ORC_* */
    int32_t use_fast;        /* b/R >= 1e-7 */
    double pos[3];
    double alpha, beta;      /* yAxisAngle, zAxisAngle */
    double T[9], Tinv[9];    /* row-major */
} orc_coil;

/* Gauss-Legendre rule of order n (1..100), ascending nodes, src/LegendreMatrix/LegendreMatrix.h:9-10.
 * Pinned by numpy.polynomial.legendre.leggauss, by the reference's own table text, and by
 * polynomial exactness (tests/test_oracle_legendre.py). */
void orc_gauss_legendre(int n, double *x, double *w);

/* src/Coil/PrecisionArguments/PrecisionFactor.cxx:15-21 */
double orc_precision_factor(double pf);

/* Coil construction. sine_frequency <= 0 means DC. */
void orc_coil_make(orc_coil *c, double R, double a, double b, int32_t N, double I, double freq,
                   const double pos[3], double alpha, double beta);

/* Increment planning, src/Coil/PrecisionArguments/PrecisionArguments.cxx:69-185 (CPU: use_gpu=0;
 * GPU: use_gpu=1 with cap gpu_increments). out = {angBlocks, thkBlocks, lenBlocks, angInc, thkInc, lenInc}. */
void orc_coil_precision_args(const orc_coil *c, double pf, int use_gpu, int gpu_increments, int32_t out[6]);
/* PrecisionArguments.cxx:188-288 */
void orc_secondary_precision_args_gpu(const orc_coil *c, double pf, int gpu_increments, int32_t out[6]);
/* CoilPairArguments.cxx:49-255; compute_gpu = (computeMethod == GPU). out[0..5] primary, out[6..11] secondary. */
void orc_pair_precision_args(const orc_coil *prim, const orc_coil *sec, double pf, int compute_gpu,
                             int z_axis_case, int pure_gpu, int gpu_increments, int32_t out[12]);

/* Per-point fields of one coil with explicit PrecisionArguments (CPU semantics: blocks, 3-D slow
 * quadrature), src/Coil/Fields/CalculateFieldsMethods.cxx:6-84, CalculateFieldsFast.cxx:13-266,
 * CalculateFieldsSlow.cxx:13-265. Any output pointer may be NULL. Points are AoS xyz.
 * G is T*G_local (the reference's convention, CalculateFieldsMethods.cxx:83), row-major 9 per point.
 * axis_ratio: on-axis branch threshold for G (CPU 1e-12, CalculateFieldsMethods.cxx:60). */
void orc_coil_fields(const orc_coil *c, const int32_t args[6], int64_t n, const double *xyz,
                     double *A, double *B, double *G, double axis_ratio);

/* Same, multi-threaded with the reference's CPU_MT chunking (CalculateAllFieldsMT.cxx:144-159);
 * quantity mask bit0 = A, bit1 = B, bit2 = G. Used as the timed CPU baseline. */
void orc_coil_fields_mt(const orc_coil *c, const int32_t args[6], int64_t n, const double *xyz,
                        double *A, double *B, double *G, double axis_ratio, int threads);

/* Same quadrature rule evaluated in x87 long double (64-bit significand): the accuracy truth used to
 * measure the error of every double and float evaluation, including the reference's own FP64 integrands
 * whose gradient terms lose up to ~1e-5 relative to cancellation near the winding. threads >= 1. */
void orc_coil_fields_ext(const orc_coil *c, const int32_t args[6], int64_t n, const double *xyz,
                         double *A, double *B, double *G, double axis_ratio, int threads);

/* Chunk boundaries of CalculateAllFieldsMT.cxx:144-159 (ends has threads+1 entries). */
void orc_chunk_ends(int64_t op_count, int threads, int64_t *ends);

/* General mutual inductance and force/torque with explicit CoilPairArguments (CPU semantics),
 * src/Coil/MInductance/CalculateMInductanceGeneral.cxx:8-117 and
 * src/Coil/ForceAndTorque/CalculateForceGeneral.cxx:8-131 (improvedPrecisionCase from
 * src/Coil/Utils/SupportFunctions.cxx:124-161, ring nodes :72-122). */
double orc_mutual_inductance_general(const orc_coil *prim, const orc_coil *sec, const int32_t args[12]);
void orc_force_torque_general(const orc_coil *prim, const orc_coil *sec, const int32_t args[12],
                              double F[3], double Tq[3]);

/* Pure-GPU pair semantics of src/Coil/CUDAKernels/MInductance.cu:24-157 and Force.cu:25-178 in FP64:
 * secondary nodes phi = pi(1+x), length b/2*x, thickness R+a/2(1+x), one block per layer.
 * correction = 2*pi*N2/I1 for M; for F/T the caller passes the correction factor it wants
 * (reference arrangement path: 2*pi*N2/I1, CalculateForceArrangements.cxx:125; physical: 2*pi*N2*I2). */
double orc_mutual_inductance_gpu_rule(const orc_coil *prim, const orc_coil *sec, const int32_t prim_inc[3],
                                      const int32_t sec_inc[3]);
void orc_force_torque_gpu_rule(const orc_coil *prim, const orc_coil *sec, const int32_t prim_inc[3],
                               const int32_t sec_inc[3], double correction, double F[3], double Tq[3]);

/* Coaxial rules (used when Coil::isZAxisCase holds, SupportFunctions.cxx:55-67): mode 0 = mutual inductance
 * (CalculateMInductanceZAxis.cxx:84-221), 1 = axial force on the secondary (CalculateForceZAxis.cxx:82-219),
 * 2 = self-inductance of prim (CalculateMInductanceZAxis.cxx:224-317; sec must be prim, the radial rule is the
 * secondary thickness arguments). extended != 0 evaluates the same rule in long double.
 * Pinned by: the general rule (orc_mutual_inductance_general / orc_force_torque_general) on coaxial pairs,
 * F = I1 I2 dM/dz by central differences, and Nagaoka's closed form of a thin solenoid's self-inductance
 * (tests/test_oracle_zaxis.py). */
double orc_zaxis_rule(const orc_coil *prim, const orc_coil *sec, const int32_t args[12], int mode, int extended);

#ifdef __cplusplus
}
#endif
#endif
