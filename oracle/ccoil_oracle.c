/*
 * ccoil_oracle.c -- FP64 CPU restatement of C-Coil's algorithms (see ccoil_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the correctness reference of the tests and the timed CPU baseline. The product never links it.
 * Built twice by oracle/Makefile: liboracle.so (IEEE, the correctness oracle) and
 * liboracle_fast.so (-O3 -mavx2 -ffast-math, the reference's own CPU flags from CMakeLists.txt:32,
 * used only as the timed CPU baseline).
 */
#include "ccoil_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
#define ORC_MU_REDUCED 1e-7                      /* g_MiReduced, CalculateFieldsFast.cxx:10 */
#define ORC_THIN_RATIO 1e-7                      /* g_thinCoilApproximationRatio, PrecisionGlobalVars.h:19 */
#define ORC_MAX_ORDER 100                        /* Legendre::maxLegendreOrder, LegendreMatrix.h:9 */
#define ORC_DEFAULT_ORDER 20                     /* PrecisionArguments.cxx:13 */
#define ORC_BASE_INC 10                          /* g_baseLayerIncrementsCPU/GPU, PrecisionGlobalVars.h:13-14 */
#define ORC_PRIM_MOD 1.414213562                 /* g_prim*WeightModifier, PrecisionGlobalVars.h:16-17 */
#define ORC_SEC_MOD 1.0                          /* g_sec*WeightModifier, PrecisionGlobalVars.h:19-20 */
#define ORC_COMMON_PLANE 1e-14                   /* g_commonPlaneApproximationRatio, PrecisionGlobalVars.h:22 */

/* ------------------------------------------------------------------------------------------ */
/* Gauss-Legendre rules (cached).                                                               */
/* ------------------------------------------------------------------------------------------ */

static double g_x[ORC_MAX_ORDER + 1][ORC_MAX_ORDER];
static double g_w[ORC_MAX_ORDER + 1][ORC_MAX_ORDER];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

/* Newton iteration on P_n from the Tricomi-type initial guess; weights 2/((1-x^2) P_n'(x)^2).
 * Nodes are symmetrised (x[n-1-i] = -x[i]) and sorted ascending like the reference table. */
static void gl_compute(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(ORC_PI * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-17) break;
        }
        /* final derivative at the converged node */
        {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            dp = (n == 1) ? 1.0 : n * (z * p1 - p0) / (z * z - 1.0);
        }
        double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        /* z is the i-th largest node */
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

static void gl_init(void)
{
    for (int n = 1; n <= ORC_MAX_ORDER; ++n) gl_compute(n, g_x[n], g_w[n]);
}

void orc_gauss_legendre(int n, double *x, double *w)
{
    pthread_once(&g_once, gl_init);
    if (n < 1 || n > ORC_MAX_ORDER) return;
    memcpy(x, g_x[n], sizeof(double) * n);
    memcpy(w, g_w[n], sizeof(double) * n);
}

static inline const double *glx(int n) { return g_x[n]; }
static inline const double *glw(int n) { return g_w[n]; }

/* ------------------------------------------------------------------------------------------ */
/* Coil construction and geometry.                                                             */
/* ------------------------------------------------------------------------------------------ */

double orc_precision_factor(double pf)
{
    /* PrecisionFactor.cxx:15-21 */
    if (pf < 1.0 || pf > 15.0) return 5.0;
    return pf;
}

static void transformation_matrices(double alpha, double beta, double T[9], double Ti[9])
{
    /* CalculateAttributes.cxx:63-78 */
    double cosY = cos(alpha), sinY = sin(alpha);
    double cosZ = cos(beta), sinZ = sin(beta);
    double t[9] = {cosZ * cosZ * cosY - sinZ * sinZ, -sinZ * cosZ * cosY - sinZ * cosZ, cosZ * sinY,
                   sinZ * cosZ * cosY + sinZ * cosZ, cosZ * cosZ - sinZ * sinZ * cosY, sinZ * sinY,
                   -sinY * cosZ, sinY * sinZ, cosY};
    double ti[9] = {cosZ * cosZ * cosY - sinZ * sinZ, sinZ * cosZ * cosY + sinZ * cosZ, -cosZ * sinY,
                    -sinZ * cosZ * cosY - sinZ * cosZ, cosZ * cosZ - sinZ * sinZ * cosY, sinZ * sinY,
                    sinY * cosZ, sinY * sinZ, cosY};
    memcpy(T, t, sizeof t);
    memcpy(Ti, ti, sizeof ti);
}

void orc_coil_make(orc_coil *c, double R, double a, double b, int32_t N, double I, double freq,
                   const double pos[3], double alpha, double beta)
{
    pthread_once(&g_once, gl_init);
    memset(c, 0, sizeof *c);
    /* Coil.cxx:45-53 (thin clamps and useFastMethod from the unclamped length) */
    c->R = R;
    c->a = (a / R < ORC_THIN_RATIO) ? 1e-18 * R * ORC_THIN_RATIO : a;
    c->b = (b / R < ORC_THIN_RATIO) ? 1e-18 * R * ORC_THIN_RATIO : b;
    c->N = N;
    c->use_fast = (b / R >= ORC_THIN_RATIO);
    /* CalculateAttributes.cxx:50-61 */
    if (c->a / R < ORC_THIN_RATIO && c->b / R < ORC_THIN_RATIO) c->type = ORC_FILAMENT;
    else if (c->a / R < ORC_THIN_RATIO) c->type = ORC_THIN;
    else if (c->b / R < ORC_THIN_RATIO) c->type = ORC_FLAT;
    else c->type = ORC_RECTANGULAR;
    /* Coil.cxx:171-175 setCurrent */
    c->I = I;
    c->J = I * N / (c->b * c->a);
    c->freq = freq > 0.0 ? freq : 0.0;
    if (pos) { c->pos[0] = pos[0]; c->pos[1] = pos[1]; c->pos[2] = pos[2]; }
    c->alpha = alpha;
    c->beta = beta;
    transformation_matrices(alpha, beta, c->T, c->Tinv);
}

/* ------------------------------------------------------------------------------------------ */
/* Increment planning (bit-exact integer restatement).                                         */
/* ------------------------------------------------------------------------------------------ */

typedef struct { int first; double second; } comp_t;
typedef struct { int first; int second; } cnt_t;

/* CoilPairArguments.cxx:20-46 */
static int balance_increments(int total, const comp_t *comps, int n, cnt_t *out)
{
    if (n < 2) {
        out[0].first = comps[0].first;
        out[0].second = total;
        return 1;
    }
    double eff = 1.0;
    for (int i = 0; i < n; ++i) eff *= comps[i].second;
    eff /= (double)total;
    double lin = pow(eff, 1.0 / (double)n);
    for (int i = 0; i < n; ++i) {
        out[i].first = comps[i].first;
        out[i].second = (int)round(comps[i].second / lin);
    }
    return n;
}

/* The do/while spill loop shared by PrecisionArguments.cxx:119-141, :237-257 and
 * CoilPairArguments.cxx:170-192. spill_mask(first) says whether a layer may be capped. */
static int plan_spill(int total, comp_t *comps, int ncomp, int use_gpu, int cap, int pure_gpu,
                      int layer_is_pair, int *layer, cnt_t *counts)
{
    int ncnt = 0;
    int spill;
    do {
        ncnt = balance_increments(total, comps, ncomp, counts);
        spill = 0;
        if (use_gpu) {
            for (int i = 0; i < ncnt; ++i) {
                int allowed = layer_is_pair ? (counts[i].first < 3 || pure_gpu) : 1;
                if (counts[i].second > cap && allowed) {
                    layer[counts[i].first] = cap;
                    for (int k = i; k < ncomp - 1; ++k) comps[k] = comps[k + 1];
                    --ncomp;
                    for (int k = i; k < ncnt - 1; ++k) counts[k] = counts[k + 1];
                    --ncnt;
                    total = (int)ceil((double)total / (double)cap);
                    spill = 1;
                    break;
                }
            }
        }
    } while (spill && ncomp > 0);
    for (int i = 0; i < ncnt; ++i) layer[counts[i].first] = counts[i].second;
    return ncnt;
}

/* PrecisionArguments.cxx:29-54 (the validating constructor) */
static void precision_arguments_ctor(int ab, int tb, int lb, int ai, int ti, int li, int32_t out[6])
{
    out[0] = ab; out[1] = tb; out[2] = lb; out[3] = ai; out[4] = ti; out[5] = li;
    if (ai > ORC_MAX_ORDER || ai < 1) out[3] = ORC_DEFAULT_ORDER;
    if (ti > ORC_MAX_ORDER || ti < 1) out[4] = ORC_DEFAULT_ORDER;
    if (li > ORC_MAX_ORDER || li < 1) out[5] = ORC_DEFAULT_ORDER;
    if (ab < 1) out[0] = 1;
    if (tb < 1) out[1] = 1;
    if (li < 1) out[2] = 1; /* the reference tests lengthIncrements here (PrecisionArguments.cxx:52) */
}

static void blocks_and_increments(const int *layer, int nl, int *blocks, int *incs)
{
    /* PrecisionArguments.cxx:166-174 / CoilPairArguments.cxx:225-233 */
    for (int i = 0; i < nl; ++i) {
        blocks[i] = (layer[i] - 1) / ORC_MAX_ORDER + 1;
        incs[i] = blocks[i] > 1 ? layer[i] / blocks[i] + 1 : layer[i];
    }
}

void orc_coil_precision_args(const orc_coil *c, double pf, int use_gpu, int cap, int32_t out[6])
{
    pf = orc_precision_factor(pf);
    int layer[3] = {1, 1, 1};
    double m[3];
    m[0] = ORC_PRIM_MOD * sqrt(ORC_PI * (c->R + 0.5 * c->a));
    m[1] = ORC_PRIM_MOD * sqrt(c->a);
    m[2] = ORC_PRIM_MOD * sqrt(c->b);
    comp_t comps[3];
    int ncomp = 0;
    int total = 1;
    comps[ncomp++] = (comp_t){0, m[0]};
    total *= ORC_BASE_INC;
    if (c->type == ORC_RECTANGULAR || c->type == ORC_FLAT) {
        total *= ORC_BASE_INC;
        comps[ncomp++] = (comp_t){1, m[1]};
    }
    total = (int)((double)total * pow(2, pf - 1.0));
    cnt_t counts[3];
    plan_spill(total, comps, ncomp, use_gpu, cap, 0, 0, layer, counts);
    if (layer[2] == 1) layer[2] = (int)round((double)layer[0] * m[2] / m[0]);
    if (layer[2] == 0) layer[2] = 1;
    if (use_gpu && layer[2] > cap) layer[2] = cap;
    int blocks[3], incs[3];
    blocks_and_increments(layer, 3, blocks, incs);
    precision_arguments_ctor(blocks[0], blocks[1], blocks[2], incs[0], incs[1], incs[2], out);
}

void orc_secondary_precision_args_gpu(const orc_coil *c, double pf, int cap, int32_t out[6])
{
    pf = orc_precision_factor(pf);
    int layer[3] = {1, 1, 1};
    double m[3];
    m[0] = ORC_PRIM_MOD * sqrt(ORC_PI * (c->R + 0.5 * c->a));
    m[1] = ORC_PRIM_MOD * sqrt(c->a);
    m[2] = ORC_PRIM_MOD * sqrt(c->b);
    comp_t comps[3];
    int ncomp = 0;
    int total = 1;
    comps[ncomp++] = (comp_t){0, m[0]};
    total *= ORC_BASE_INC;
    switch (c->type) {
        case ORC_RECTANGULAR:
            total *= ORC_BASE_INC * ORC_BASE_INC;
            comps[ncomp++] = (comp_t){1, m[1]};
            comps[ncomp++] = (comp_t){2, m[2]};
            break;
        case ORC_THIN:
            total *= ORC_BASE_INC;
            comps[ncomp++] = (comp_t){2, m[2]};
            break;
        case ORC_FLAT:
            total *= ORC_BASE_INC;
            comps[ncomp++] = (comp_t){1, m[1]};
            break;
        default:
            break;
    }
    total = (int)((double)total * pow(2, pf - 1.0));
    cnt_t counts[3];
    plan_spill(total, comps, ncomp, 1, cap, 0, 0, layer, counts);
    if (layer[2] == 1) layer[2] = (int)round((double)layer[0] * m[2] / m[0]);
    if (layer[2] == 0) layer[2] = 1;
    if (layer[2] > cap) layer[2] = cap;
    precision_arguments_ctor(1, 1, 1, layer[0], layer[1], layer[2], out);
}

void orc_pair_precision_args(const orc_coil *prim, const orc_coil *sec, double pf, int compute_gpu,
                             int z_axis_case, int pure_gpu, int cap, int32_t out[12])
{
    pf = orc_precision_factor(pf);
    int layer[6] = {1, 1, 1, 1, 1, 1};
    double m[6];
    m[0] = ORC_PRIM_MOD * sqrt(ORC_PI * (prim->R + 0.5 * prim->a));
    m[1] = ORC_PRIM_MOD * sqrt(prim->a);
    m[2] = ORC_PRIM_MOD * sqrt(prim->b);
    m[3] = ORC_SEC_MOD * sqrt(2 * ORC_PI * (sec->R + 0.5 * sec->a));
    m[4] = ORC_SEC_MOD * sqrt(sec->a);
    m[5] = ORC_SEC_MOD * sqrt(sec->b);
    const int pb = ORC_BASE_INC, sb = ORC_BASE_INC;
    comp_t comps[6];
    int ncomp = 0;
    int total = 1;
    comps[ncomp++] = (comp_t){0, m[0]};
    total *= pb;
    if (prim->type == ORC_RECTANGULAR || prim->type == ORC_FLAT) {
        total *= pb;
        comps[ncomp++] = (comp_t){1, m[1]};
    }
    if (z_axis_case) {
        if (sec->type == ORC_RECTANGULAR || sec->type == ORC_FLAT) {
            total *= sb;
            comps[ncomp++] = (comp_t){4, m[4]};
        }
    } else {
        comps[ncomp++] = (comp_t){3, m[3]};
        total *= sb;
        switch (sec->type) {
            case ORC_RECTANGULAR:
                total *= sb * sb;
                comps[ncomp++] = (comp_t){4, m[4]};
                comps[ncomp++] = (comp_t){5, m[5]};
                break;
            case ORC_THIN:
                total *= sb;
                comps[ncomp++] = (comp_t){5, m[5]};
                break;
            case ORC_FLAT:
                total *= sb;
                comps[ncomp++] = (comp_t){4, m[4]};
                break;
            default:
                break;
        }
    }
    total = (int)((double)total * pow(2, pf - 1.0));
    cnt_t counts[6];
    plan_spill(total, comps, ncomp, compute_gpu, cap, pure_gpu, 1, layer, counts);
    for (int i = 1; i < 6; ++i) {
        if (layer[i] == 1) layer[i] = (int)round((double)layer[0] * m[i] / m[0]);
        if (layer[i] == 0) layer[i] = 1;
        if (compute_gpu && layer[i] >= cap) layer[i] = cap;
    }
    int blocks[6], incs[6];
    blocks_and_increments(layer, 6, blocks, incs);
    precision_arguments_ctor(blocks[0], blocks[1], blocks[2], incs[0], incs[1], incs[2], out);
    precision_arguments_ctor(blocks[3], blocks[4], blocks[5], incs[3], incs[4], incs[5], out + 6);
}

/* ------------------------------------------------------------------------------------------ */
/* Local (cylindrical) field integrals, CPU semantics, in double and in long double.            */
/* ------------------------------------------------------------------------------------------ */

#define REAL double
#define SUF(name) name##_d
#define SQRT sqrt
#define LOG log
#define COS cos
#define SIN sin
#define ATAN2 atan2
#include "ccoil_oracle_fields.inc"
#undef REAL
#undef SUF
#undef SQRT
#undef LOG
#undef COS
#undef SIN
#undef ATAN2

#define REAL long double
#define SUF(name) name##_ld
#define SQRT sqrtl
#define LOG logl
#define COS cosl
#define SIN sinl
#define ATAN2 atan2l
#include "ccoil_oracle_fields.inc"
#undef REAL
#undef SUF
#undef SQRT
#undef LOG
#undef COS
#undef SIN
#undef ATAN2

/* extended != 0 selects the long double evaluation */
static void point_fields(const orc_coil *c, const int32_t args[6], const double p[3], double *A, double *B,
                         double *G, double axis_ratio, int extended)
{
    if (extended) point_fields_ld(c, args, p, A, B, G, axis_ratio);
    else point_fields_d(c, args, p, A, B, G, axis_ratio);
}

void orc_coil_fields(const orc_coil *c, const int32_t args[6], int64_t n, const double *xyz,
                     double *A, double *B, double *G, double axis_ratio)
{
    pthread_once(&g_once, gl_init);
    for (int64_t i = 0; i < n; ++i)
        point_fields(c, args, xyz + 3 * i, A ? A + 3 * i : NULL, B ? B + 3 * i : NULL,
                     G ? G + 9 * i : NULL, axis_ratio, 0);
}

void orc_chunk_ends(int64_t op_count, int threads, int64_t *ends)
{
    /* CalculateAllFieldsMT.cxx:144-159 */
    int64_t average = (int64_t)floor((double)op_count / (double)threads);
    int64_t remaining = op_count;
    ends[0] = 0;
    for (int i = 0; i < threads; ++i) {
        int64_t temp = (remaining % (threads - i) == 0) ? average : average + 1;
        ends[i + 1] = (op_count - remaining) + temp;
        remaining -= temp;
    }
}

typedef struct {
    const orc_coil *c;
    const int32_t *args;
    const double *xyz;
    double *A, *B, *G;
    double axis_ratio;
    int64_t start, stop;
    int extended;
} mt_task;

static void *mt_run(void *p)
{
    mt_task *t = (mt_task *)p;
    for (int64_t i = t->start; i < t->stop; ++i)
        point_fields(t->c, t->args, t->xyz + 3 * i, t->A ? t->A + 3 * i : NULL, t->B ? t->B + 3 * i : NULL,
                     t->G ? t->G + 9 * i : NULL, t->axis_ratio, t->extended);
    return NULL;
}

static void fields_mt(const orc_coil *c, const int32_t args[6], int64_t n, const double *xyz, double *A,
                      double *B, double *G, double axis_ratio, int threads, int extended)
{
    pthread_once(&g_once, gl_init);
    if (threads < 1) threads = 1;
    int64_t *ends = (int64_t *)malloc(sizeof(int64_t) * (threads + 1));
    orc_chunk_ends(n, threads, ends);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    mt_task *tasks = (mt_task *)malloc(sizeof(mt_task) * threads);
    for (int i = 0; i < threads; ++i) {
        tasks[i] = (mt_task){c, args, xyz, A, B, G, axis_ratio, ends[i], ends[i + 1], extended};
        pthread_create(&th[i], NULL, mt_run, &tasks[i]);
    }
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(tasks);
    free(th);
    free(ends);
}

void orc_coil_fields_mt(const orc_coil *c, const int32_t args[6], int64_t n, const double *xyz,
                        double *A, double *B, double *G, double axis_ratio, int threads)
{
    pthread_once(&g_once, gl_init);
    fields_mt(c, args, n, xyz, A, B, G, axis_ratio, threads, 0);
}

void orc_coil_fields_ext(const orc_coil *c, const int32_t args[6], int64_t n, const double *xyz,
                         double *A, double *B, double *G, double axis_ratio, int threads)
{
    pthread_once(&g_once, gl_init);
    fields_mt(c, args, n, xyz, A, B, G, axis_ratio, threads < 1 ? 1 : threads, 1);
}

/* ------------------------------------------------------------------------------------------ */
/* Pair quantities.                                                                            */
/* ------------------------------------------------------------------------------------------ */

static void cross(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* SupportFunctions.cxx:124-161 */
static int improved_precision_case(const orc_coil *prim, const orc_coil *sec, double *offset)
{
    double pn[3] = {sin(prim->alpha) * cos(prim->beta), sin(prim->alpha) * sin(prim->beta), cos(prim->alpha)};
    double sn[3] = {sin(sec->alpha) * cos(sec->beta), sin(sec->alpha) * sin(sec->beta), cos(sec->alpha)};
    double diff[3] = {prim->pos[0] - sec->pos[0], prim->pos[1] - sec->pos[1], prim->pos[2] - sec->pos[2]};
    double plane[3];
    cross(pn, diff, plane);
    double dot = plane[0] * sn[0] + plane[1] * sn[1] + plane[2] * sn[2];
    int improved = fabs(dot) <= ORC_COMMON_PLANE;
    *offset = 0.0;
    if (improved) {
        if (plane[0] == 0.0 && plane[1] == 0.0 && plane[2] == 0.0) cross(sn, diff, plane);
        if (plane[0] == 0.0 && plane[1] == 0.0 && plane[2] == 0.0) cross(sn, pn, plane);
        *offset = atan2(cos(sec->alpha) * (plane[0] * cos(sec->beta) + plane[1] * sin(sec->beta)) -
                            plane[2] * sin(sec->alpha),
                        plane[0] * sin(sec->beta) - plane[1] * cos(sec->beta));
    }
    return improved;
}

/* SupportFunctions.cxx:72-122: unit ring position and tangent at each angular node. */
static void ring_nodes(int ablocks, int ainc, const orc_coil *sec, int improved, double offset,
                       double *posv, double *tanv)
{
    double blk = 2 * ORC_PI / ablocks;
    if (improved) blk = ORC_PI / ablocks;
    double sA = sin(sec->alpha), cA = cos(sec->alpha), sB = sin(sec->beta), cB = cos(sec->beta);
    int k = 0;
    for (int pb = 0; pb < ablocks; ++pb) {
        double bp = offset + blk * (pb + 0.5);
        for (int pi = 0; pi < ainc; ++pi, ++k) {
            double phi = bp + (blk * 0.5) * glx(ainc)[pi];
            double sP = sin(phi), cP = cos(phi);
            posv[3 * k + 0] = cB * cA * cP - sB * sP;
            posv[3 * k + 1] = sB * cA * cP + cB * sP;
            posv[3 * k + 2] = (-1) * sA * cP;
            tanv[3 * k + 0] = (-1) * cB * cA * sP - sB * cP;
            tanv[3 * k + 1] = (-1) * sB * cA * sP + cB * cP;
            tanv[3 * k + 2] = sA * sP;
        }
    }
}

/* Secondary node cloud of CalculateMInductanceGeneral.cxx:41-103 / CalculateForceGeneral.cxx:41-103. */
static int64_t secondary_cloud(const orc_coil *sec, const int32_t sargs[6], int improved, double offset,
                               double **pts, double **wts, double **tan_unit, int *ring_count)
{
    int lb = sargs[2], li = sargs[5], tb = sargs[1], ti = sargs[4], ab = sargs[0], ai = sargs[3];
    int64_t n = (int64_t)lb * li * tb * ti * ab * ai;
    int rc = ab * ai;
    double *rp = (double *)malloc(sizeof(double) * 3 * rc);
    double *rt = (double *)malloc(sizeof(double) * 3 * rc);
    ring_nodes(ab, ai, sec, improved, offset, rp, rt);
    double *P = (double *)malloc(sizeof(double) * 3 * n);
    double *W = (double *)malloc(sizeof(double) * n);
    double lbs = sec->b / lb, tbs = sec->a / tb;
    double sA = sin(sec->alpha), cA = cos(sec->alpha), sB = sin(sec->beta), cB = cos(sec->beta);
    int64_t k = 0;
    for (int zb = 0; zb < lb; ++zb)
        for (int rb = 0; rb < tb; ++rb) {
            double zbp = (-1) * (sec->b * 0.5) + lbs * (zb + 0.5);
            double rbp = sec->R + tbs * (rb + 0.5);
            for (int zi = 0; zi < li; ++zi)
                for (int ri = 0; ri < ti; ++ri) {
                    double ringR = rbp + (tbs * 0.5) * glx(ti)[ri];
                    double ld = zbp + (lbs * 0.5) * glx(li)[zi];
                    for (int pb = 0; pb < ab; ++pb)
                        for (int pi = 0; pi < ai; ++pi, ++k) {
                            int pp = pb * ai + pi;
                            P[3 * k + 0] = sec->pos[0] + ld * sA * cB + ringR * rp[3 * pp + 0];
                            P[3 * k + 1] = sec->pos[1] + ld * sA * sB + ringR * rp[3 * pp + 1];
                            P[3 * k + 2] = sec->pos[2] + ld * cA + ringR * rp[3 * pp + 2];
                            W[k] = 0.125 * ringR * glw(li)[zi] * glw(ti)[ri] * glw(ai)[pi];
                        }
                }
        }
    free(rp);
    *pts = P;
    *wts = W;
    *tan_unit = rt;
    *ring_count = rc;
    return n;
}

double orc_mutual_inductance_general(const orc_coil *prim, const orc_coil *sec, const int32_t args[12])
{
    pthread_once(&g_once, gl_init);
    double offset;
    int improved = improved_precision_case(prim, sec, &offset);
    double *P, *W, *tan_u;
    int rc;
    int64_t n = secondary_cloud(sec, args + 6, improved, offset, &P, &W, &tan_u, &rc);
    double *A = (double *)malloc(sizeof(double) * 3 * n);
    orc_coil_fields(prim, args, n, P, A, NULL, NULL, 1e-12);
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        int p = (int)(i % rc);
        m += (A[3 * i] * tan_u[3 * p] + A[3 * i + 1] * tan_u[3 * p + 1] + A[3 * i + 2] * tan_u[3 * p + 2]) * W[i];
    }
    m /= (args[6 + 2] * args[6 + 1] * args[6 + 0]);
    free(A); free(P); free(W); free(tan_u);
    return m * 2 * ORC_PI * sec->N / prim->I;
}

void orc_force_torque_general(const orc_coil *prim, const orc_coil *sec, const int32_t args[12],
                              double F[3], double Tq[3])
{
    pthread_once(&g_once, gl_init);
    double *P, *W, *tan_u;
    int rc;
    int64_t n = secondary_cloud(sec, args + 6, 0, 0.0, &P, &W, &tan_u, &rc);
    double *B = (double *)malloc(sizeof(double) * 3 * n);
    orc_coil_fields(prim, args, n, P, NULL, B, NULL, 1e-12);
    double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        int p = (int)(i % rc);
        double tf[3];
        cross(tan_u + 3 * p, B + 3 * i, tf);
        tf[0] *= W[i]; tf[1] *= W[i]; tf[2] *= W[i];
        f[0] += tf[0]; f[1] += tf[1]; f[2] += tf[2];
        double arm[3] = {P[3 * i] - sec->pos[0], P[3 * i + 1] - sec->pos[1], P[3 * i + 2] - sec->pos[2]};
        double tt[3];
        cross(arm, tf, tt);
        t[0] += tt[0]; t[1] += tt[1]; t[2] += tt[2];
    }
    double factor = 2 * ORC_PI * (sec->I * sec->N) / (args[6 + 2] * args[6 + 1] * args[6 + 0]);
    for (int k = 0; k < 3; ++k) { F[k] = f[k] * factor; Tq[k] = t[k] * factor; }
    free(B); free(P); free(W); free(tan_u);
}

/* Pure-GPU rule (MInductance.cu:24-157, Force.cu:25-178) restated in FP64. */
static void gpu_rule_pair(const orc_coil *prim, const orc_coil *sec, const int32_t prim_inc[3],
                          const int32_t sec_inc[3], int want_force, double correction, double *M,
                          double F[3], double Tq[3])
{
    int32_t pargs[6] = {1, 1, 1, prim_inc[0], prim_inc[1], prim_inc[2]};
    int la = sec_inc[0], lt = sec_inc[1], ll = sec_inc[2];
    double sA = sin(sec->alpha), cA = cos(sec->alpha), sB = sin(sec->beta), cB = cos(sec->beta);
    double m = 0.0, f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    for (int ia = 0; ia < la; ++ia)
        for (int it = 0; it < lt; ++it)
            for (int il = 0; il < ll; ++il) {
                double lp = sec->b * 0.5 * glx(ll)[il];
                double tp = sec->R + sec->a * 0.5 * (1.0 + glx(lt)[it]);
                double ap = ORC_PI * (1.0 + glx(la)[ia]);
                double sP = sin(ap), cP = cos(ap);
                double rpos[3] = {cB * cA * cP - sB * sP, sB * cA * cP + cB * sP, -sA * cP};
                double rtan[3] = {-cB * cA * sP - sB * cP, -sB * cA * sP + cB * cP, sA * sP};
                double ring[3] = {lp * sA * cB + tp * rpos[0], lp * sA * sB + tp * rpos[1], lp * cA + tp * rpos[2]};
                double p[3] = {sec->pos[0] + ring[0], sec->pos[1] + ring[1], sec->pos[2] + ring[2]};
                double w = 0.125 * tp * glw(ll)[il] * glw(lt)[it] * glw(la)[ia];
                if (!want_force) {
                    double A[3];
                    point_fields(prim, pargs, p, A, NULL, NULL, 1e-12, 0);
                    m += w * (A[0] * rtan[0] + A[1] * rtan[1] + A[2] * rtan[2]);
                } else {
                    double B[3], df[3], dt[3];
                    point_fields(prim, pargs, p, NULL, B, NULL, 1e-12, 0);
                    cross(rtan, B, df);
                    for (int k = 0; k < 3; ++k) df[k] *= correction * w;
                    cross(ring, df, dt);
                    for (int k = 0; k < 3; ++k) { f[k] += df[k]; t[k] += dt[k]; }
                }
            }
    if (!want_force) *M = m * 2 * ORC_PI * sec->N / prim->I;
    else for (int k = 0; k < 3; ++k) { F[k] = f[k]; Tq[k] = t[k]; }
}

double orc_mutual_inductance_gpu_rule(const orc_coil *prim, const orc_coil *sec, const int32_t prim_inc[3],
                                      const int32_t sec_inc[3])
{
    pthread_once(&g_once, gl_init);
    double m = 0.0;
    gpu_rule_pair(prim, sec, prim_inc, sec_inc, 0, 0.0, &m, NULL, NULL);
    return m;
}

void orc_force_torque_gpu_rule(const orc_coil *prim, const orc_coil *sec, const int32_t prim_inc[3],
                               const int32_t sec_inc[3], double correction, double F[3], double Tq[3])
{
    pthread_once(&g_once, gl_init);
    gpu_rule_pair(prim, sec, prim_inc, sec_inc, 1, correction, NULL, F, Tq);
}

#define REAL double
#define SUF(name) name##_d
#define SQRT sqrt
#define LOG log
#define COS cos
#include "ccoil_oracle_zaxis.inc"
#undef REAL
#undef SUF
#undef SQRT
#undef LOG
#undef COS

#define REAL long double
#define SUF(name) name##_ld
#define SQRT sqrtl
#define LOG logl
#define COS cosl
#include "ccoil_oracle_zaxis.inc"
#undef REAL
#undef SUF
#undef SQRT
#undef LOG
#undef COS

double orc_zaxis_rule(const orc_coil *prim, const orc_coil *sec, const int32_t args[12], int mode, int extended)
{
    pthread_once(&g_once, gl_init);
    if (mode < 0 || mode > 2) return 0.0;
    if (extended) return (double)zaxis_rule_ld(prim, sec, args, mode);
    return zaxis_rule_d(prim, sec, args, mode);
}
