/*
 * ccoil_hip.h -- C ABI of the MI355X (gfx950) compute path for C-Coil.
 *
 * This is the drop-in boundary that replaces C-Coil's CUDA kernel interface:
 *   - Calculate_hardware_accelerated_{a,b,g}        src/Coil/CUDAKernels/CoilAcceleration.h:8-18
 *   - Calculate_hardware_accelerated_{a,b,e,g}_group src/CoilGroup/CUDAKernels/CoilGroupAcceleration.h:7-21
 *   - Calculate_{mutual_inductance,force_and_torque}_configurations
 *                                                   src/Coil/CUDAKernels/CoilAcceleration.h:21-29
 *   - Calculate_{mutual_inductance,force_and_torque}_configurations_group
 *                                                   src/CoilGroup/CUDAKernels/CoilGroupAcceleration.h:23-37
 * Differences from that interface, all deliberate:
 *   - C linkage, plain pointers and sizes, status codes (the reference exits the process on any
 *     device error, src/CUDAUtils/ErrorCheck/ErrorCheck.cu:5-17);
 *   - a context owns the device, stream and buffers (the reference keeps process globals,
 *     src/CUDAUtils/MemoryManagement/GPUMemoryManagement.cu:10-13);
 *   - coils are passed as SoA FP64 physical attributes; the library derives the per-coil constants
 *     (constFactor, transformation matrices) exactly as generateCoilData does
 *     (src/Coil/Utils/SupportFunctions.cxx:167-223) and generates the Gauss-Legendre tables on the device;
 *   - one call evaluates all coils and every requested quantity (A, B, E, dB) in one launch.
 *
 * Threading: a context may be used by one thread at a time. Calls on one context are ordered on its stream.
 */
#ifndef CCOIL_HIP_H
#define CCOIL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccoil_ctx_s *ccoil_ctx;
typedef int32_t ccoil_status;

enum {
    CCOIL_OK = 0,
    CCOIL_ERR_ARGUMENT = -1,      /* invalid sizes, pointers or precision values; nothing was launched */
    CCOIL_ERR_OUT_OF_MEMORY = -2, /* device or host allocation failed (reference: std::bad_alloc) */
    CCOIL_ERR_DEVICE = -3,        /* a HIP runtime call failed; see ccoil_last_error */
    CCOIL_ERR_NO_DEVICE = -4,     /* no HIP device available (reference: "GPU functions are disabled") */
    CCOIL_ERR_CAPACITY = -5       /* a plan's launch needed more FP32C near-pass records than the plan holds */
};

/* Quantity mask for ccoil_fields. */
enum {
    CCOIL_A = 1,          /* vector potential A */
    CCOIL_B = 2,          /* flux density B */
    CCOIL_E = 4,          /* E = 2*pi*f*A per coil (ComputeFieldsMethods.cxx:27-32, FieldEGroup.cu:105-107) */
    CCOIL_G = 8,          /* gradient in the reference convention T*G_local (CalculateFieldsMethods.cxx:83) */
    CCOIL_G_JACOBIAN = 16 /* true Jacobian dB_i/dx_j = T*G_local*T^T (no reference counterpart) */
};

/* Arithmetic of the quadrature. */
enum {
    CCOIL_FP64 = 0,           /* cancellation-free integrands in double; reciprocals and reciprocal square roots
                                 from the hardware FP64 approximations refined by one Newton step (a few ulp):
                                 the accurate mode */
    CCOIL_FP32C = 1,          /* the same integrands with FP32 per-node arithmetic, FP64 per-row work and row
                                 accumulation, re-evaluation with CCOIL_FP64's arithmetic of near-winding (coil,
                                 point) pairs and, for outputs that divide by r (the gradient, the pair position
                                 gradient, the adjoint), of pairs near the coil's axis */
    CCOIL_FP32 = 2,           /* the reference's integrands in float (what a TYPE=float build computes) */
    CCOIL_FP64_REFERENCE = 3, /* the reference's integrands verbatim in double (a TYPE=double build): the
                                 parity mode and the plain-FP64 throughput comparator; its gradient terms
                                 lose up to ~3e-5 relative to cancellation near the winding */
    CCOIL_FP64_IEEE = 4       /* CCOIL_FP64's integrands with IEEE division and square root */
};

/* Flags. */
enum {
    CCOIL_DEVICE_POINTERS = 1, /* points/outputs are device pointers on the context's device */
    CCOIL_NO_NEAR_FP64 = 2,    /* FP32C: do not re-evaluate near-winding or near-axis pairs in FP64 */
    CCOIL_FORCE_REFERENCE_SCALE = 4, /* ccoil_pairs: force/torque scaled by 2*pi*N2/I1 as the reference's
                                       single-pair arrangement path does (CalculateForceArrangements.cxx:125,
                                       SupportFunctions.cxx:238-241); default is the physical 2*pi*N2*I2 */
    CCOIL_AUTOTUNE = 8,        /* ccoil_fields, ccoil_pairs and the plans: take the launch shape (field: lanes per
                                  point and coil chunks; pairs: primary-node evaluations per lane and work item)
                                  from the context's tuning cache; the first call of a size not in the cache times
                                  every candidate shape on the device (the call runs once per candidate) and caches
                                  the fastest. A cached shape is replayed as is, so results are bitwise
                                  reproducible for a given cache. */
    CCOIL_DIRECT = 16,         /* FP32C fields: evaluate every (coil, point) pair with the rule. By default a pair
                                  beyond the coil's far-field distance (at least twice the radius of the sphere that
                                  holds the winding, more for coils of few angular increments) takes the coil's
                                  exterior zonal-harmonic series in FP64 instead, with moments summed over the rule's
                                  own thickness nodes: it departs from the FP64 rule by less than 1e-11 of the coil's
                                  field there (the angular rule's error bound) and from the float rule it replaces
                                  by the float rule's own rounding. No counterpart in the reference. */
    CCOIL_FAR_FIELD = 32       /* CCOIL_FP64 fields: take the far-field series as CCOIL_FP32C does, with the
                                  threshold at 1e-14 of the coil's field (the FP64 rule's own rounding level). No
                                  counterpart in the reference; without it CCOIL_FP64 evaluates every pair with
                                  the rule. */
};

/* Pair quantity mask for ccoil_pairs. CCOIL_PAIR_DM: gradient of M with respect to the secondary's position,
 * the exact derivative of the same quadrature (no counterpart in the reference; F / (I1 I2) approximates it up to
 * the quadrature error, and dM/d(primary position) = -dM). */
enum { CCOIL_PAIR_M = 1, CCOIL_PAIR_FT = 2, CCOIL_PAIR_DM = 4 };

/* Coils, SoA, host memory, FP64. Lengths are n. Increments are the already-planned quadrature orders
 * (1..100) from PrecisionArguments; blocks apply to pair secondaries only. */
typedef struct {
    int64_t n;
    const int32_t *use_fast;   /* 1: THIN/RECTANGULAR (analytic length integral), 0: FLAT/FILAMENT */
    const int32_t *n_ang;      /* angular increments */
    const int32_t *n_thk;      /* thickness increments */
    const int32_t *n_len;      /* length increments (pair secondaries; NULL for fields) */
    const int32_t *blocks;     /* 3 per coil: angular, thickness, length blocks (pair secondaries; NULL = 1) */
    const double *R;           /* inner radius */
    const double *a;           /* thickness after the thin clamp (Coil.cxx:27-28) */
    const double *b;           /* length after the thin clamp */
    const double *J;           /* current density N*I/(a*b) */
    const double *current;     /* I */
    const double *turns;       /* N */
    const double *frequency;   /* sine frequency for E (NULL: E not requested) */
    const double *pos;         /* 3 per coil */
    const double *alpha;       /* rotation about y (yAxisAngle) */
    const double *beta;        /* rotation about z (zAxisAngle) */
} ccoil_coils;

/* Per-call report (any pointer argument may be NULL). */
typedef struct {
    int64_t node_evaluations;  /* sum over (coil, point) of quadrature nodes, or (pair) node-pairs */
    int64_t near_pairs;        /* FP32C: (coil, point) pairs re-evaluated in FP64 */
    double kernel_ms;          /* device time of the compute kernels of this call (hipEvent) */
    int32_t block_size;        /* threads per workgroup of the main kernel */
    int32_t split;             /* fields: lanes cooperating on one point; pairs: most work items one pair's
                                  primary thickness rows were split over */
    double main_kernel_ms;     /* device time of the main kernel alone (field_kernel or pair_kernel; hipEvent) */
    double record_ms;          /* fields with a record pass: device time of its evaluation kernels (hipEvent) */
} ccoil_report;

/* Context lifecycle. ccoil_last_error(NULL) returns the last error of ccoil_create on this thread. */
ccoil_status ccoil_create(int device, ccoil_ctx *out);
ccoil_status ccoil_destroy(ccoil_ctx ctx);
const char *ccoil_last_error(ccoil_ctx ctx);
const char *ccoil_version(void);
/* Use an external stream (hipStream_t cast to void*; hipStreamLegacy binds the default null stream); NULL restores
 * the context's own (non-blocking) stream. */
ccoil_status ccoil_set_stream(ccoil_ctx ctx, void *stream);
ccoil_status ccoil_synchronize(ccoil_ctx ctx);
/* Autotuner cache file (no counterpart in the reference): loads the entries of path (a text file, one "key<TAB>shape"
 * line per tuned call size; missing file: empty) into the context's cache and appends every new entry to it. NULL
 * keeps the cache in memory only. A context also loads the file named by the environment variable
 * CCOIL_TUNING_FILE when it is created. */
ccoil_status ccoil_set_tuning_file(ccoil_ctx ctx, const char *path);
/* Far-field series geometry (no counterpart in the reference; host only, no context or device): for each coil, the
 * radius r_max of the sphere about its centre that holds its winding and the distance beyond which a field call of
 * the given precision takes the coil's series (CCOIL_FP32C by default, CCOIL_FP64 with CCOIL_FAR_FIELD), as pair
 * calls and field calls with one lane per point (16,384 points or more in these modes) use it; field calls with fewer
 * points take the series from at least 2 r_max. 0: the coil never does (few angular increments). Other precisions
 * give 0. Errors: ccoil_last_error(NULL). */
ccoil_status ccoil_far_field_radius(const ccoil_coils *coils, int32_t precision, double *rmax, double *distance);
/* The stream the context's kernels are ordered on (its own, or the one given to ccoil_set_stream), as hipStream_t
 * cast to void*. No counterpart in the reference: zero-copy callers hand it to a producer (DLPack __dlpack__(stream))
 * so that the producer's pending writes are ordered before the context's kernels. */
ccoil_status ccoil_get_stream(ccoil_ctx ctx, void **stream);

/* Gauss-Legendre rule of order n (1..100) as generated in FP64 on the device; replaces the host table
 * src/LegendreMatrix/LegendreMatrix.h:9-10. x, w: host arrays of n doubles, ascending nodes. */
ccoil_status ccoil_gauss_legendre(ccoil_ctx ctx, int32_t n, double *x, double *w);

/* Fields of all coils summed at n_points points (AoS xyz). Outputs AoS: A, B, E are 3 per point, G is 9
 * per point row-major; each requested output is overwritten. Replaces the per-coil launch loops of
 * FieldBGroup.cu:142-208, FieldGGroup.cu:212-278, FieldAGroup.cu, FieldEGroup.cu and the single-coil
 * FieldA/B/G.cu launchers (a one-coil ccoil_coils). */
ccoil_status ccoil_fields(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                          uint32_t quantities, int32_t precision, uint32_t flags,
                          double *A, double *B, double *E, double *G, ccoil_report *report);

/* ccoil_fields that allocates nothing. No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Every buffer a field
 * call needs beyond its inputs and outputs (coil records, coil-chunk partials, the FP32C near pass) comes from a
 * fixed 64 MiB scratch array per device, held from the library's load until the process ends. Points and outputs
 * must be device pointers (flags must hold CCOIL_DEVICE_POINTERS). A call whose buffers do not fit, or whose FP32C
 * near pass finds more near pairs than the scratch holds, returns CCOIL_ERR_ARGUMENT (in the second case the
 * outputs lack the near contributions): split the points. Results are bitwise those of ccoil_fields. All calls on
 * one device share its scratch, which keeps the last call's contents: calls that overlap on one device (from two
 * contexts or threads) race. */
ccoil_status ccoil_fields_noalloc(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                                  uint32_t quantities, int32_t precision, uint32_t flags,
                                  double *A, double *B, double *E, double *G, ccoil_report *report);

/* Mutual inductance M and force/torque (F, T about the secondary centre) for n_pairs (primary, secondary)
 * pairs, each pair in its own pose. prim_index/sec_index select coils of prim/sec (NULL: identity).
 * ring_offset/half_ring (per pair, NULL: 0) reproduce the coplanar improved-precision rule of
 * Coil::improvedPrecisionCase (SupportFunctions.cxx:124-161). M is 1 per pair, F, T and dM 3 per pair.
 * Replaces MInductance.cu:184-248, Force.cu:206-270 and the *Group.cu launchers. */
ccoil_status ccoil_pairs(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                         const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                         const double *ring_offset, const int32_t *half_ring,
                         uint32_t what, int32_t precision, uint32_t flags,
                         double *M, double *F, double *T, double *dM, ccoil_report *report);

/* Gradient of each pair's mutual inductance with respect to both coils' parameters: the adjoint of M (no counterpart
 * in the reference, whose force and torque are its physical counterparts: F = I1 I2 dM/d(secondary position),
 * CalculateForceGeneral.cxx:114-128, and the torque about the secondary's centre is I1 I2 times the derivative of M
 * with respect to a rotation of the secondary). Same pairs, rule and scale as ccoil_pairs with CCOIL_PAIR_M; the
 * derivatives are those of that quadrature, with every quadrature order and the coplanar half-ring rule held fixed.
 * grad: 16 per pair, dM/d(secondary x, y, z, alpha, beta, R, a, b) then dM/d(primary x, y, z, alpha, beta, R, a, b)
 * (alpha: yAxisAngle, beta: zAxisAngle; a, b: the thickness and length as given). M does not depend on either coil's
 * current. M (nullable): 1 per pair. precision: CCOIL_FP64, CCOIL_FP64_IEEE or CCOIL_FP32C. */
ccoil_status ccoil_pairs_grad(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                              const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                              const double *ring_offset, const int32_t *half_ring, int32_t precision, uint32_t flags,
                              double *M, double *grad, ccoil_report *report);

/* Adjoint of the summed B field (no counterpart in the reference's CUDA interface; it backs the gradient of
 * scalar objectives of the field, e.g. a torch.autograd.Function over ccoil_fields). Given dL_dB (3 per point,
 * global frame) for a scalar L of B = sum_c B_c at n_points points, writes per coil c the 9 values
 *   dL/d(current, R, a, b, px, py, pz, alpha, beta)   (grad, 9 per coil, coil-major)
 * of sum_x dL_dB(x) . B_c(x), with each coil's quadrature orders (n_ang, n_thk) held fixed. Needs coil current
 * and turns (B is linear in the current through J = N I / (a b)). precision: CCOIL_FP64, CCOIL_FP64_IEEE or
 * CCOIL_FP32C (near-winding and near-axis pairs in FP64). Deterministic: partial sums are reduced in a fixed order. */
ccoil_status ccoil_field_gradient(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                                  const double *dL_dB, int32_t precision, uint32_t flags, double *grad,
                                  ccoil_report *report);

/* Coaxial pairs (the z-axis case of Coil::isZAxisCase, SupportFunctions.cxx:55-67: both axes on the global z axis)
 * with the analytic length integral of both coils, for THIN and RECTANGULAR coils (use_fast = 1). For each pair
 * writes M (what & CCOIL_PAIR_M) and Fz, the axial force on the secondary (what & CCOIL_PAIR_FT; the transverse
 * force and the torque of a coaxial pair are zero). Only pos[2] of each coil enters. Primary rule: n_ang, n_thk
 * and the angular and thickness blocks; secondary rule: n_thk and the thickness blocks (blocks NULL = 1). Needs
 * J, current (M divides by I1 I2) and turns. Evaluated in FP64 with the reference's per-node expressions.
 * Replaces Coil::calculateMutualInductanceZAxisFast (CalculateMInductanceZAxis.cxx:84-221) and
 * Coil::calculateAmpereForceZAxisFast (CalculateForceZAxis.cxx:82-219). */
ccoil_status ccoil_pairs_zaxis(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                               const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                               uint32_t what, uint32_t flags, double *M, double *Fz, ccoil_report *report);

/* Self-inductance of n = coils->n THIN or RECTANGULAR coils (L, 1 per coil). Coil rule: n_ang, n_thk and the
 * angular and thickness blocks; radial: 2 per coil (thickness blocks, thickness increments) of the secondary
 * precision the reference plans for the pair (coil, coil). Needs J, current and turns.
 * Replaces Coil::calculateSelfInductance (CalculateMInductanceZAxis.cxx:224-317). */
ccoil_status ccoil_self_inductance(ccoil_ctx ctx, const ccoil_coils *coils, const int32_t *radial, uint32_t flags,
                                   double *L, ccoil_report *report);

/* Plans (no counterpart in the reference): one field or pair call recorded once and captured into a HIP graph that
 * every launch replays with one submission, no host work and no host synchronisation. The coil records, the pair
 * work list, the launch shape (with CCOIL_AUTOTUNE, the tuned one) and every buffer are fixed when the plan is
 * made; points and outputs must be device pointers (CCOIL_DEVICE_POINTERS), and their contents may change between
 * launches while their addresses may not. Making a plan runs the call once (FP32C field plans: to size the near
 * pass, which then holds a quarter more (coil, point) pairs than that call found). The context must outlive its
 * plans, and a plan's launches must not overlap one another. */
typedef struct ccoil_plan_s *ccoil_plan;
ccoil_status ccoil_fields_plan(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                               uint32_t quantities, int32_t precision, uint32_t flags,
                               double *A, double *B, double *E, double *G, ccoil_plan *plan);
ccoil_status ccoil_pairs_plan(ccoil_ctx ctx, const ccoil_coils *prim, const ccoil_coils *sec,
                              const int64_t *prim_index, const int64_t *sec_index, int64_t n_pairs,
                              const double *ring_offset, const int32_t *half_ring, uint32_t what, int32_t precision,
                              uint32_t flags, double *M, double *F, double *T, double *dM, ccoil_plan *plan);
/* Enqueues the plan's graph on stream (a hipStream_t cast to void*; NULL: the context's stream) and returns. */
ccoil_status ccoil_plan_launch(ccoil_plan plan, void *stream);
/* Waits for the plan's last launch and checks it: CCOIL_ERR_CAPACITY when an FP32C field launch found more near
 * pairs than the plan holds (its outputs then lack the near contributions; make a new plan). report: the device
 * time of the whole launch (kernel_ms = main_kernel_ms) and the launch's node evaluations and near pairs. */
ccoil_status ccoil_plan_synchronize(ccoil_plan plan, ccoil_report *report);
ccoil_status ccoil_plan_destroy(ccoil_plan plan);

#ifdef __cplusplus
}
#endif
#endif
