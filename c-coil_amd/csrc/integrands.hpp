// Quadrature integrands of one (coil, point) pair in the coil's local cylindrical frame.
//
// eval_plain<T>: the reference's integrands verbatim in arithmetic T (double: the FP64 comparator and
//   the near-winding evaluator; float: what a TYPE=float build of the reference computes).
//   Fast (THIN/RECTANGULAR) B: FieldBGroup.cu:47-77, G: FieldGGroup.cu:257-305, A: FieldAGroup.cu;
//   slow (FLAT/FILAMENT) B: FieldBGroup.cu:78-101, G: FieldGGroup.cu:306-339, A: FieldAGroup.cu.
// eval_stable<T>: the same integrands, algebraically rearranged so that no subtraction cancels; per-node
//   work in T (float: FP32C; double: the accurate FP64 mode), per-row work and row sums in FP64.
//
// Both evaluate thickness rows j = row0, row0 + rstep, ... so several lanes can share one point.
#pragma once

#include "common.hpp"

namespace ccoil {

template <typename T> __device__ inline T rsq(T x);
template <> __device__ inline double rsq<double>(double x) { return 1.0 / sqrt(x); }
template <> __device__ inline float rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }

template <typename T> __device__ inline T div_(T a, T b);
template <> __device__ inline double div_<double>(double a, double b) { return a / b; }
template <> __device__ inline float div_<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <typename T> __device__ inline T sqrt_(T x);
template <> __device__ inline double sqrt_<double>(double x) { return sqrt(x); }
template <> __device__ inline float sqrt_<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T> __device__ inline T log_(T x);
template <> __device__ inline double log_<double>(double x) { return log(x); }
template <> __device__ inline float log_<float>(float x) { return __logf(x); }

template <typename T> __device__ inline const T *tab_x(const Tables &t);
template <> __device__ inline const double *tab_x<double>(const Tables &t) { return t.x; }
template <> __device__ inline const float *tab_x<float>(const Tables &t) { return t.xf; }
template <typename T> __device__ inline const T *tab_w(const Tables &t);
template <> __device__ inline const double *tab_w<double>(const Tables &t) { return t.w; }
template <> __device__ inline const float *tab_w<float>(const Tables &t) { return t.wf; }
template <typename T> __device__ inline const T *tab_cos(const Tables &t);
template <> __device__ inline const double *tab_cos<double>(const Tables &t) { return t.cosv; }
template <> __device__ inline const float *tab_cos<float>(const Tables &t) { return t.cosf; }

template <typename T, bool WA, bool WB, bool WG>
__device__ inline void eval_plain(const CoilDev &c, const Tables &tab, double zd, double rd, int row0, int rstep,
                                  Local &L)
{
    const T *xt = tab_x<T>(tab) + table_offset(c.n_thk);
    const T *wt = tab_w<T>(tab) + table_offset(c.n_thk);
    const T *ct = tab_cos<T>(tab) + table_offset(c.n_ang);
    const T *wa = tab_w<T>(tab) + table_offset(c.n_ang);
    const int na = c.n_ang, nt = c.n_thk;
    const T cf = (T)c.cf, R = (T)c.R, a = (T)c.a, b = (T)c.b;
    const T z = (T)zd, r = (T)rd;
    T sA = 0, sBh = 0, sBz = 0, sRP = 0, sRR = 0, sRZ = 0, sZZ = 0;

    if (c.use_fast) {
        const T top = z + (T)0.5 * b, bot = z - (T)0.5 * b;
        for (int j = row0; j < nt; j += rstep) {
            const T rho = R + (T)0.5 * a * ((T)1 + xt[j]);
            const T wtj = wt[j];
            // B: FieldBGroup.cu:51-58
            const T A_ = rho * rho;
            const T Bb = (T)2 * rho * r;
            const T C_ = A_ + r * r;
            const T D1 = top * top + C_, D2 = bot * bot + C_;
            // G: FieldGGroup.cu:261-272
            const T gB = r * r, gC = rho * r, gD = A_ + gB, gE = A_ * A_ + gB * gB, gF = gC * gC;
            // A: FieldAGroup.cu (tempConstA/B)
            const T aA = A_ + r * r, aB = (T)2 * rho * r;
#pragma unroll 4
            for (int i = 0; i < na; ++i) {
                const T cp = ct[i];
                const T g = cf * wtj * wa[i];
                if (WB) {
                    const T E_ = Bb * cp;
                    const T F1 = rsq<T>(D1 - E_), F2 = rsq<T>(D2 - E_);
                    sBh += g * rho * cp * (F2 - F1);
                    sBz += g * div_<T>(A_ - (T)0.5 * E_, C_ - E_) * (top * F1 - bot * F2);
                }
                if (WG) {
                    const T cp2 = cp * cp;
                    const T phiE = (T)2 * gC * cp;
                    const T J1 = gD + top * top - phiE, J2 = gD + bot * bot - phiE;
                    const T K1 = rsq<T>(J1), K2 = rsq<T>(J2);
                    const T L1 = div_<T>(K1, J1), L2 = div_<T>(K2, J2);
                    const T M_ = gD - phiE;
                    const T N_ = (T)2 * gF * cp * (cp2 + (T)2) - gC * ((T)3 * cp2 + (T)1) * gD + cp * gE;
                    const T O_ = cp * gD - (T)2 * gC;
                    sRP += g * div_<T>(rho * cp, r) * (K2 - K1);
                    sRR += g * (gC - A_ * cp) * cp * (L1 - L2);
                    sZZ += g * (A_ - gC * cp) * (L1 - L2);
                    sRZ += g * div_<T>(rho, M_ * M_) * (top * L1 * (O_ * J1 + N_) - bot * L2 * (O_ * J2 + N_));
                }
                if (WA) {
                    const T C = rsq<T>(aA - aB * cp);
                    const T d1 = top * C, d2 = bot * C;
                    const T e1 = sqrt_<T>(d1 * d1 + (T)1), e2 = sqrt_<T>(d2 * d2 + (T)1);
                    sA += g * rho * cp * log_<T>(div_<T>(e1 + d1, e2 + d2));
                }
            }
        }
    } else {
        for (int j = row0; j < nt; j += rstep) {
            const T rho = R + (T)0.5 * a * ((T)1 + xt[j]);
            const T wtj = wt[j];
            // B: FieldBGroup.cu:82-87
            const T A_ = rho * rho, B_ = rho * r, C_ = A_ + r * r + z * z, D_ = rho * z;
            // G: FieldGGroup.cu:310-319
            const T gB = r * r, gC = z * z, gD = r * rho, gE = rho * z;
            const T gF = (T)2 * A_ + (T)2 * gB - gC, gG = A_ + gB + gC;
#pragma unroll 4
            for (int i = 0; i < na; ++i) {
                const T cp = ct[i];
                const T h = cf * wtj * wa[i];
                if (WB) {
                    const T E_ = C_ - (T)2 * B_ * cp;
                    const T F_ = E_ * sqrt_<T>(E_);
                    const T G_ = div_<T>(h, F_);
                    sBh += G_ * (D_ * cp);
                    sBz += G_ * (A_ - B_ * cp);
                }
                if (WG) {
                    const T I_ = gG - (T)2 * gD * cp;
                    const T J_ = I_ * sqrt_<T>(I_);
                    const T X = div_<T>(h, J_), Y = div_<T>(h, J_ * I_);
                    sRP += div_<T>(X * gE * cp, r);
                    sRR += Y * ((T)-3 * gE * (r - rho * cp)) * cp;
                    sZZ += Y * ((T)-3 * gE * (rho - r * cp));
                    sRZ += Y * (rho * (gF - gD * cp) * cp - (T)3 * A_ * r);
                }
                if (WA) {
                    const T aA = A_ + r * r + z * z, aB = (T)2 * rho * r;
                    sA += h * rho * cp * rsq<T>(aA - aB * cp);
                }
            }
        }
    }
    L.A = (double)sA;
    L.Bh = (double)sBh;
    L.Bz = (double)sBz;
    L.RP = (double)sRP;
    L.RR = (double)sRR;
    L.RZ = (double)sRZ;
    L.ZZ = (double)sZZ;
}

// ---------------------------------------------------------------------------------------------
// Cancellation-free integrands (the same discretisation as eval_plain, algebraically rearranged).
// Notation: rho = source radius of the row, d = rho - r, base = d^2, q = 2 rho r,
// s = 1 - cos(phi) (table), M = base + q s (in-plane distance squared: a sum of non-negative terms),
// J_k = M + z_k^2 with z_1 = z + b/2, z_2 = z - b/2, K_k = J_k^-1/2. The fast rule is evaluated at z >= 0 (the
// sheet is mirror-symmetric; odd terms flip sign afterwards), so z1 > 0 and z2 < 0 exactly inside the axial span.
//   K1 - K2       = -2 b z (K1 K2)^2 / (K1 + K2)                  (exact identity, no cancellation)
//   z1 K1 - z2 K2 = 2 b z M (K1 K2)^2 / (z1 K1 + z2 K2)            when z2 > 0
//                 = z1 K1 + |z2| K2 (direct)                        otherwise
//   z1 K1^3 - z2 K2^3 = (z1 K1 - z2 K2) K1^2 - 2 b z z2 K2 (K1 K2)^2  (both terms non-negative when z2 < 0)
//   K1^3 - K2^3   = (K1 - K2) ((K1 + K2)^2 - K1 K2)
//   N (gradient RZ term) as a polynomial in s with coefficients in base and q.
// One reciprocal serves 1/(K1 + K2), 1/M and 1/(z1 K1 + z2 K2): with Y = M (z1 K1 + z2 K2) (or M) and
// Rb = 1 / ((K1 + K2) Y), 1/(K1 + K2) = Y Rb and 1/Y = (K1 + K2) Rb. The angular loops are unrolled by 4 so
// that the scalar table loads of several nodes are in flight together.
// The factor rho common to every integrand and the thickness weight are applied per row in FP64, and row
// sums are accumulated in FP64, so T = float computes only the per-node work in FP32 (FP32C).
// ---------------------------------------------------------------------------------------------
template <typename T> __device__ inline T rcp_(T x);
template <> __device__ inline double rcp_<double>(double x) { return 1.0 / x; }
template <> __device__ inline float rcp_<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T> __device__ inline T fma_(T a, T b, T c);
template <> __device__ inline double fma_<double>(double a, double b, double c) { return fma(a, b, c); }
template <> __device__ inline float fma_<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T> __device__ inline T log1p_(T x);
template <> __device__ inline double log1p_<double>(double x) { return log1p(x); }
template <> __device__ inline float log1p_<float>(float x) { return log1pf(x); }

// Arithmetic policies of eval_stable: float (FP32C), double (IEEE division and square root), and f64s:
// double arithmetic whose reciprocals and reciprocal square roots come from the hardware FP64
// approximations (v_rcp_f64, v_rsq_f64; relative error below 2^-24 on gfx950) refined by one Newton step.
// The refined value errs by about (3/8) e^2 for the reciprocal square root and e^2 for the reciprocal, i.e.
// below 1e-15 relative, so f64s is accurate to a few ulp per operation while avoiding the scaled division and
// square-root sequences (FP64 FMA issues at the FP32 rate on CDNA3/4; those sequences dominate the IEEE
// double kernel's time). Its domain is that of the approximations: normal, finite operands.
struct f64s {};
template <typename P> struct Real { using type = P; };
template <> struct Real<f64s> { using type = double; };
template <typename P> using real_t = typename Real<P>::type;

template <typename P> __device__ inline real_t<P> prsq(real_t<P> x) { return rsq<real_t<P>>(x); }
template <> __device__ inline double prsq<f64s>(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);               // 1 - x y^2
    return fma(0.5 * y, e, y);                            // y (1 + e/2)
}
template <typename P> __device__ inline real_t<P> prcp(real_t<P> x) { return rcp_<real_t<P>>(x); }
template <> __device__ inline double prcp<f64s>(double x)
{
    const double y = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, y, 1.0);                      // 1 - x y
    return fma(y, e, y);                                   // y (1 + e)
}

template <typename T> __device__ inline T frexp_(T x, int *e);
template <> __device__ inline double frexp_<double>(double x, int *e) { return frexp(x, e); }
template <> __device__ inline float frexp_<float>(float x, int *e) { return frexpf(x, e); }

// 2 atanh(u) / u as a polynomial in t = u^2 on [0, (3 - 2 sqrt 2)^2] (the fit log1p_pos describes).
template <typename T> __device__ inline T atanh2_poly(T t)
{
    T s;
    if constexpr (sizeof(T) == 8) {
        s = (T)0.14809710360655276;
        s = fma_<T>(s, t, (T)0.1531252814836419);
        s = fma_<T>(s, t, (T)0.18183631680229329);
        s = fma_<T>(s, t, (T)0.22222197056726048);
        s = fma_<T>(s, t, (T)0.2857142876064168);
        s = fma_<T>(s, t, (T)0.3999999999930234);
        s = fma_<T>(s, t, (T)0.6666666666666765);
        s = fma_<T>(s, t, (T)2.0);
    } else {
        s = (T)0.2992439f;
        s = fma_<T>(s, t, (T)0.3997485f);
        s = fma_<T>(s, t, (T)0.6666682f);
        s = fma_<T>(s, t, (T)2.0f);
    }
    return s;
}

// log(1 + x) for x >= 0 in one branch-free instruction stream (the lanes of a wave take different x):
// 1 + x = m 2^e with m in [sqrt(1/2), sqrt(2)), log(1 + x) = e ln 2 + 2 atanh(u), u = (m - 1) / (m + 1), |u| <= 3 - 2 sqrt 2.
// 2 atanh(u) / u is a polynomial in t = u^2 on [0, (3 - 2 sqrt 2)^2]: a Chebyshev fit of degree 7 (double; relative
// error 3e-18 with the rounded coefficients below) or 3 (float; 1.7e-9), in place of the Taylor series of atanh.
// For e = 0 the numerator is x itself (exact), so small x keep full relative accuracy; for e != 0, m - 1 is exact
// (Sterbenz) and log(1 + x) >= 0.34 bounds the effect of rounding 1 + x.
template <typename P> __device__ inline real_t<P> log1p_pos(real_t<P> x)
{
    using T = real_t<P>;
    int e;
    T m = frexp_<T>((T)1 + x, &e);                       // m in [1/2, 1)
    const bool lo = m < (T)0.70710678118654752440;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const T num = e == 0 ? x : m - (T)1;
    const T u = num * prcp<P>(m + (T)1);
    return fma_<T>((T)e, (T)0.69314718055994530942, u * atanh2_poly<T>(u * u));
}

// log(1 + x) for x = N / D >= 0, given v = x / (2 + x) = N / (2 D + N) from a reciprocal the caller shares with other
// terms: 2 atanh(v) = v atanh2_poly(v^2) while v <= 3 - 2 sqrt 2 (x <= sqrt 2 - 1: log1p_pos's e = 0 case, without its
// range reduction and its own reciprocal); lanes beyond that (nodes close to the winding) take log1p_pos(N / D), with
// D = den() computed only there. A lane's value depends only on its own arguments.
constexpr double kLog1pRatioMax = 0.17157287525380990;   // 3 - 2 sqrt 2
template <typename P, typename F> __device__ inline real_t<P> log1p_ratio(real_t<P> N, real_t<P> v, F den)
{
    using T = real_t<P>;
    T r = v * atanh2_poly<T>(v * v);
    if (v > (T)kLog1pRatioMax) r = log1p_pos<P>(N * prcp<P>(den()));
    return r;
}

template <typename T> __device__ inline const T *tab_wcos(const Tables &t);
template <> __device__ inline const double *tab_wcos<double>(const Tables &t) { return t.wcos; }
template <> __device__ inline const float *tab_wcos<float>(const Tables &t) { return t.wcosf; }

template <typename T> __device__ inline const T *tab_opc(const Tables &t);
template <> __device__ inline const double *tab_opc<double>(const Tables &t) { return t.opc; }
template <> __device__ inline const float *tab_opc<float>(const Tables &t) { return t.opcf; }

template <typename T> __device__ inline const T *tab_wcs(const Tables &t);
template <> __device__ inline const double *tab_wcs<double>(const Tables &t) { return t.wcs; }
template <> __device__ inline const float *tab_wcs<float>(const Tables &t) { return t.wcsf; }

template <typename T> __device__ inline const T *tab_omc(const Tables &t);
template <> __device__ inline const double *tab_omc<double>(const Tables &t) { return t.omc; }
template <> __device__ inline const float *tab_omc<float>(const Tables &t) { return t.omcf; }

// Adjoint extras of one (coil, point): derivatives of the local (B_rho, B_z) with respect to the coil's inner
// radius R, thickness a and length b ([0] rho component, [1] z component). Each quadrature row j is the field
// f_j(rho_j, b, r, z) of a current sheet (fast method) or ring (slow method) of radius rho_j = R + a/2 (1 + x_j);
// f_j is homogeneous in (rho, b, r, z) of degree 0 (sheet) or -1 (ring), so
//   rho df/drho = -(deg f + r df/dr + z df/dz + b df/db),
// with df/dr, df/dz the row's own gradient terms and df/db = (ring at z + b/2 + ring at z - b/2) / 2. The
// constant factors hold J a = N I / b (sheet: dB/db gains -B/b) and J a b = N I (ring: no a or b).
struct AdjLocal {
    double dR[2], da[2], db[2];
};

// Pair adjoint extras of one (coil, point): the radial derivative of the local A_phi of the same rule, dA/dr (the
// discrete B_z equals A/r + dA/dr only up to the angular rule's error, so the derivative is summed node by node), and
// the derivatives of A_phi with respect to the coil's inner radius R, thickness a and length b. Row j of the rule is
// A_j = W_j S_j with W_j = cf w_j rho_j: for the sheet (fast method) cf = 1e-7 N I pi / (2 b) and
// S_j = sum w cos (asinh(z1 / sqrt M) - asinh(z2 / sqrt M)), whose node terms have dS/dM = -(z1 K1 - z2 K2) / (2 M)
// and dM/dr = 2 (r - rho cos), so dA_j/dr = -W_j sum w cos (r - rho cos) (z1 K1 - z2 K2) / M; A_j is homogeneous of
// degree 0 in (rho, b, r, z). For the ring (slow method) cf = 1e-7 N I pi / 2, S_j = sum w cos K, dA_j/dr =
// -W_j sum w cos (r - rho cos) K^3, and A_j is homogeneous of degree 0 in (rho, r, z). Hence
//   rho dA_j/drho = -(r dA_j/dr + z dA_j/dz + b dA_j/db),  dA_j/dz = -B_rho,j (exact node by node),
// and, for the sheet, dA_j/db = -A_j / b + W_j sum w cos (K1 + K2) / 2 (the end rings).
struct AdjA {
    double dR, da, db, Ar;
};

// Angular-loop unrolling of eval_stable: 4, except 2 for the float B / gradient instances (the FP32C field kernel
// measured 2 % faster with 2 and 12 % slower with 8).
template <typename T, bool WA> constexpr int kStableUnroll = (sizeof(T) == 4 && !WA) ? 2 : 4;

template <typename P, bool WA, bool WB, bool WG, bool AD = false, bool ADA = false, bool LEAN = false>
__device__ inline void eval_stable(const CoilDev &c, const Tables &tab, double zd, double rd, int row0, int rstep,
                                   Local &L, AdjLocal *ad = nullptr, AdjA *ada = nullptr)
{
    static_assert(!AD || (WB && WG), "the adjoint needs B and the gradient terms");
    static_assert(!ADA || (WA && WB), "the pair adjoint needs A and B");
    using T = real_t<P>;
    const double *xt = tab.x + table_offset(c.n_thk);
    const double *wt = tab.w + table_offset(c.n_thk);
    const T *st = tab_omc<T>(tab) + table_offset(c.n_ang);
    const T *wa = tab_w<T>(tab) + table_offset(c.n_ang);
    const T *wct = tab_wcos<T>(tab) + table_offset(c.n_ang);   // w cos
    const T *opt = tab_opc<T>(tab) + table_offset(c.n_ang);    // 1 + cos
    const T *wcst = tab_wcs<T>(tab) + table_offset(c.n_ang);     // w cos (1 - cos)
    const double *wsd = tab.ws + table_offset(c.n_ang);          // w (1 - cos), double fast instances without G
    const int na = c.n_ang, nt = c.n_thk;
    double aA = 0, aBh = 0, aBz = 0, aRR = 0, aRZ = 0, aZZ = 0;
    double hR0 = 0, hR1 = 0, ha0 = 0, ha1 = 0, hb0 = 0, hb1 = 0;   // adjoint row sums
    double qR = 0, qa = 0, qb = 0, qAr = 0;                         // pair adjoint row sums (A_phi)

    if (c.use_fast) {
        // The sheet's field is mirror-symmetric in z: B_rho, dB_rho/drho, dB_z/dz (and the adjoint's rho
        // components) are odd, A, B_z and dB_rho/dz even. The rule is evaluated at |z|, so the upper edge z1 is
        // positive and z2 > 0 exactly when the point lies beyond the sheet's axial span; odd terms flip at the end.
        const bool neg = zd < 0.0;
        const double zm = fabs(zd);
        const double z1d = zm + 0.5 * c.b, z2d = zm - 0.5 * c.b;
        const T z1 = (T)z1d, z2 = (T)z2d;
        const T z1s = (T)(z1d * z1d), z2s = (T)(z2d * z2d);
        const T tbz = (T)(2.0 * c.b * zm), zs = (T)(2.0 * zm);
        const bool same = z2d > 0.0;
        // same ? x : 1 as one exact FMA, fma(sg, x, sg1): a 64-bit select takes two instructions
        const T sg = same ? (T)1 : (T)0, sg1 = (T)1 - sg;
        const T bf = (T)c.b;
        // The rows, for the wave's lanes all beyond the span (kAll: the selects above resolve at compile time) or
        // for any mix; the two instances give bitwise the same values (fma(0, x, 1) = 1 and fma(1, x, 0) = x).
        auto rows = [&](auto all_same) {
            constexpr bool kAll = decltype(all_same)::value;
            // LEAN (the pair kernels, whose waves are fixed by the pair's schedule): with every lane beyond the span
            // and no gradient terms, the node loop below takes fewer operations; its values may differ from the
            // general loop's in the last bits, so a lane's value depends on its wave there.
            constexpr bool kLean = LEAN && kAll && !WG && !AD && !ADA;
            const bool sm = kAll || same;
            for (int j = row0; j < nt; j += rstep) {
                const double rho_d = c.R + 0.5 * c.a * (1.0 + xt[j]);
                const double dd = rho_d - rd;
                const T base = (T)(dd * dd);
                const T q = (T)(2.0 * rho_d * rd);
                const T dm = (T)dd;                 // rho - r
                const T rf = (T)rd, rho = (T)rho_d;
                const T base2 = (T)(dd * dd * dd * dd);
                const T n1 = (T)(dd * dd * (2.0 * rho_d * rd - dd * dd));
                const T n2 = (T)(-3.0 * rho_d * rd * dd * dd);
                const T n3 = (T)(-2.0 * rho_d * rho_d * rd * rd);
                T rBh = 0, rBz = 0, rRZ = 0, rZZ = 0, rA = 0;
                T rRRs = 0, rRRc = 0;   // sum w cos s dL and sum w cos dL: the row's RR is rho rRRs - dm rRRc
                T rBzs = 0;             // double without the gradient: sum w s dgM (rBz then holds sum w dgM)
                // adjoint: end rings, sum w cos (z1 K1^3 + z2 K2^3), sum w (rho - r cos)(K1^3 + K2^3)
                T eR = 0, eZ = 0;
                T eA = 0;           // pair adjoint: end rings of A, sum w cos (K1 + K2)
                T rAr = 0;          // pair adjoint: sum w cos (r - rho cos) dgM
#pragma unroll kStableUnroll<T, WA>
                for (int i = 0; i < na; ++i) {
                    const T s = st[i], w = wa[i], wc = wct[i];
                    const T M = fma_<T>(q, s, base);
                    const T J1 = M + z1s, J2 = M + z2s;
                    const T K1 = prsq<P>(J1), K2 = prsq<P>(J2);
                    if constexpr (kLean) {
                        // z2 > 0: with gs = z1 K1 + z2 K2, (z1 K1 - z2 K2) / M = tbz K1^2 K2^2 / gs and
                        // K2 - K1 = tbz K1^2 K2^2 / (K1 + K2), so the B sums are taken without their common factor
                        // tbz (applied per row) and one reciprocal gives 1 / (K1 + K2), 1 / gs and 1 / Ea, where
                        // Ea = 2 Da + numr = sum (2 sJ2 + 2 |z|) + tbz (all terms positive).
                        const T K12 = K1 * K2, K12s = K12 * K12;
                        T sJa2 = 0, sum = 0, numr = 0, Re = 0;
                        if (WA) {
                            sJa2 = J2 * K2;
                            sum = fma_<T>(J1, K1, sJa2);
                            numr = fma_<T>(bf, sum, tbz);
                        }
                        const T Ea = WA ? fma_<T>(sum, fma_<T>((T)2, sJa2, zs), tbz) : (T)1;
                        if (WB) {
                            const T Ks = K1 + K2;
                            const T gs = fma_<T>(z1, K1, z2 * K2);
                            const T KG = Ks * gs;
                            const T R = prcp<P>(WA ? KG * Ea : KG);
                            const T Rb = WA ? Ea * R : R;                       // 1 / (Ks gs)
                            if (WA) Re = KG * R;
                            rBh = fma_<T>(wc, K12s * (gs * Rb), rBh);          // B_rho / tbz
                            const T dg = K12s * (Ks * Rb);                      // (z1 K1 - z2 K2) / (M tbz)
                            if constexpr (sizeof(T) == 8) {
                                rBz = fma_<T>(w, dg, rBz);
                                rBzs = fma_<T>((T)wsd[i], dg, rBzs);
                            } else {
                                rBz = fma_<T>(w * fma_<T>(rf, s, dm), dg, rBz);
                            }
                        } else if (WA) {
                            Re = prcp<P>(Ea);
                        }
                        if (WA)
                            rA = fma_<T>(wc, log1p_ratio<P>(numr, numr * Re, [&] { return sum * (z2 + sJa2); }), rA);
                        continue;
                    }
                    // A: asinh(z1 / sqrt(M)) - asinh(z2 / sqrt(M)) = log(P1 / P2), P_k = z_k + sqrt(M + z_k^2),
                    // = log1p(x) with x = (P1 - P2) / P2 = (b (sJ1 + sJ2) + 2 b |z|) / ((sJ1 + sJ2) P2) >= 0;
                    // for z2 < 0, P2 = M / (sJ2 - z2). A is even in z, so |z| serves directly. x = Nx / Da with
                    // Nx = numr * nsel, and x / (2 + x) = Nx / Ea, Ea = 2 Da + Nx (log1p_ratio).
                    T Nx = 0, Da = 1, Ea = 1;
                    if (WA) {
                        const T sJa1 = J1 * K1, sJa2 = J2 * K2;
                        const T sum = sJa1 + sJa2;
                        const T numr = fma_<T>(bf, sum, tbz);
                        const T nsel = kAll ? (T)1 : fma_<T>(sg1, sJa2 - z2, sg);   // same ? 1 : sJ2 - z2
                        Da = sum * (sm ? z2 + sJa2 : M);
                        Nx = numr * nsel;
                        Ea = fma_<T>(Da, (T)2, Nx);
                    }
                    T Re = 0;   // 1 / Ea
                    if (WB || WG) {
                        const T K12 = K1 * K2;
                        const T tP2 = tbz * (K12 * K12);
                        const T g2 = z2 * K2;
                        const T gs = fma_<T>(z1, K1, g2);                   // z1 K1 + z2 K2
                        const T gsel = kAll ? gs : fma_<T>(sg, gs, sg1);    // same ? gs : 1
                        const T Y = M * gsel;
                        const T Ks = K1 + K2;
                        const T KsY = Ks * Y;
                        // one reciprocal for 1 / (K1 + K2), 1 / Y and, with A, 1 / Ea
                        const T R = prcp<P>(WA ? KsY * Ea : KsY);
                        const T Rb = WA ? Ea * R : R;
                        if (WA) Re = KsY * R;
                        const T S = Y * Rb, Rr = Ks * Rb;
                        const T dK = -tP2 * S;                              // K1 - K2
                        const T invM = gsel * Rr;
                        const T dgM = (sm ? tP2 * M : fma_<T>(z1, K1, -g2)) * Rr;   // (z1 K1 - z2 K2) / M
                        rBh = fma_<T>(wc, -dK, rBh);       // B_rho, also the gradient's rho-phi term
                        if constexpr (ADA) {
                            eA = fma_<T>(wc, Ks, eA);
                            rAr = fma_<T>(wc * fma_<T>(rho, s, -dm), dgM, rAr);   // r - rho cos = rho s - (rho - r)
                        }
                        constexpr bool kBzMoments = sizeof(T) == 8 && !WG;
                        T wr = 0, rmc = 0;
                        if constexpr (kBzMoments) {
                            // without the gradient, B_z as dm sum w dgM + r sum w s dgM: two FMAs with tabulated
                            // weights
                            if (WB) {
                                rBz = fma_<T>(w, dgM, rBz);
                                rBzs = fma_<T>((T)wsd[i], dgM, rBzs);
                            }
                        } else {
                            rmc = fma_<T>(rf, s, dm);                           // rho - r cos
                            wr = w * rmc;                                       // shared by B_z and dB_z/dz
                            if (WB) rBz = fma_<T>(wr, dgM, rBz);
                        }
                        if (WG) {
                            const T K1s = K1 * K1;
                            // K1^3 - K2^3 = (K1 - K2)(K1^2 + K1 K2 + K2^2), the second factor as (K1 + K2)^2 - K1 K2
                            // (the subtraction removes at most a quarter: no cancellation)
                            const T dL = dK * fma_<T>(Ks, Ks, -K12);
                            // r - rho cos = rho s - dm: two moments, each one FMA with a tabulated weight; their
                            // combination per row (in double) is bounded by the same sum of magnitudes as the per-node
                            // terms
                            rRRs = fma_<T>(wcst[i], dL, rRRs);
                            rRRc = fma_<T>(wc, dL, rRRc);
                            rZZ = fma_<T>(wr, dL, rZZ);
                            // hM = (z1 K1^3 - z2 K2^3) / M = dgM K1^2 + z2 K2 (K1^2 - K2^2) / M, K1^2 - K2^2 = -tP2:
                            // with z >= 0 both terms are positive when z2 < 0, and dgM carries no cancellation when
                            // z2 > 0
                            const T hM = fma_<T>(dgM, K1s, -(tP2 * g2) * invM);
                            // cos (rho^2 + r^2) - 2 rho r = cos base - q s = (1 + cos) base - M
                            const T O = fma_<T>(opt[i], base, -M);
                            const T N = fma_<T>(s, fma_<T>(s, fma_<T>(s, n3, n2), n1), base2);
                            // the rho-z term (O (z1 K1 - z2 K2) + N (z1 K1^3 - z2 K2^3)) / M^2
                            const T u = fma_<T>(O, dgM, N * hM);
                            rRZ = fma_<T>(w, u * invM, rRZ);
                            if constexpr (AD) {
                                const T K13 = K1s * K1, K23 = K2 * K2 * K2;
                                eR = fma_<T>(wc, fma_<T>(z1, K13, z2 * K23), eR);
                                eZ = fma_<T>(w, rmc * (K13 + K23), eZ);
                            }
                        }
                    } else if (WA) {
                        Re = prcp<P>(Ea);
                    }
                    if (WA) rA = fma_<T>(wc, log1p_ratio<P>(Nx, Nx * Re, [&] { return Da; }), rA);
                }
                const double rowRR = WG ? fma(rho_d, (double)rRRs, -dd * (double)rRRc) : 0.0;
                const double rowBz = (sizeof(T) == 8 && !WG) ? fma(rd, (double)rBzs, dd * (double)rBz) : (double)rBz;
                const double W = c.cf * wt[j] * rho_d;
                const double Wb = kLean ? W * (2.0 * c.b * zm) : W;   // the lean loop's B sums lack tbz
                aBh += Wb * (double)rBh;
                aBz += Wb * rowBz;
                aRR += W * rowRR;
                aZZ += W * (double)rZZ;
                aRZ += W * (double)rRZ;
                aA += W * (double)rA;
                if constexpr (ADA) {
                    const double rowA = W * (double)rA, dAdb = W * 0.5 * (double)eA - rowA / c.b;
                    const double dAdr = -W * (double)rAr;
                    const double dAdrho = -(rd * dAdr - zm * W * (double)rBh + c.b * dAdb) / rho_d;
                    qR += dAdrho;
                    qa += 0.5 * (1.0 + xt[j]) * dAdrho;
                    qb += dAdb;
                    qAr += dAdr;
                }
                if constexpr (AD) {
                    const double Wr = c.cf * wt[j], hx = 0.5 * (1.0 + xt[j]);
                    const double e0 = 0.5 * (double)eR, e1 = 0.5 * (double)eZ;
                    const double g0 = rd * rowRR + zm * (double)rRZ + c.b * e0;         // -d f_rho / d rho
                    const double g1 = rd * (double)rRZ + zm * (double)rZZ + c.b * e1;   // -d f_z / d rho
                    hR0 += Wr * g0;
                    hR1 += Wr * g1;
                    ha0 += Wr * hx * g0;
                    ha1 += Wr * hx * g1;
                    hb0 += W * e0;
                    hb1 += W * e1;
                }
            }
        };
        if (__ballot(same) == __ballot(true)) rows(std::true_type{});
        else rows(std::false_type{});
        const double sgn = neg ? -1.0 : 1.0;   // odd terms at z < 0
        if constexpr (ADA) {   // A is even in z
            ada->dR = qR;
            ada->da = qa;
            ada->db = qb;
            ada->Ar = qAr;
        }
        if constexpr (AD) {
            ad->dR[0] = -sgn * hR0;
            ad->dR[1] = -hR1;
            ad->da[0] = -sgn * ha0;
            ad->da[1] = -ha1;
            ad->db[0] = sgn * (hb0 - aBh / c.b);
            ad->db[1] = hb1 - aBz / c.b;
        }
        L.A = aA;
        L.Bh = sgn * aBh;
        L.Bz = aBz;
        L.RP = L.Bh / rd;
        L.RR = sgn * aRR;
        L.RZ = aRZ;
        L.ZZ = sgn * aZZ;
    } else {
        for (int j = row0; j < nt; j += rstep) {
            const double rho_d = c.R + 0.5 * c.a * (1.0 + xt[j]);
            const double dd = rho_d - rd;
            const double zz = zd * zd;
            const T baseJ = (T)(dd * dd + zz);
            const T q = (T)(2.0 * rho_d * rd);
            const T dm = (T)dd;
            const T rf = (T)rd, rho = (T)rho_d;
            const T k0 = (T)(2.0 * dd * dd - zz);
            const T k1 = (T)(2.0 * dd * dd + 2.0 * rho_d * rd - zz);
            const T k2 = (T)(rho_d * rd);
            T rBh = 0, rBz = 0, rRR = 0, rRZ = 0, rZZ = 0, rA = 0;
            T sRRs = 0, sRRc = 0;   // float: sum w cos s K^5 and sum w cos K^5
            T rC2 = 0;              // pair adjoint: sum w cos^2 K^3 (float: accumulated as sum w cos s K^3 first)
            if constexpr (sizeof(T) == 8) {
                // Double: the angular sums as moments in s = 1 - cos of positive terms, b_k = sum w s^k K^3 and
                // g_k = sum w s^k K^5, combined once per row: (1 - s)(rho s - d), d + r s and the rho-z polynomial
                // are at most quadratic in s. The combinations cancel no more than the per-node sums they replace
                // (the cos-weighted first harmonic). The weights w s and w s^2 are tabulated, so each moment is one
                // FMA per node.
                const double *wst = tab.ws + table_offset(c.n_ang), *ws2t = tab.ws2 + table_offset(c.n_ang);
                T b0 = 0, b1 = 0, b2 = 0, g0 = 0, g1 = 0, g2 = 0;
#pragma unroll kStableUnroll<T, WA>
                for (int i = 0; i < na; ++i) {
                    const T s = st[i], w = wa[i];
                    const T J = fma_<T>(q, s, baseJ);
                    const T K = prsq<P>(J);
                    if (WB || WG) {
                        const T K2 = K * K;
                        const T K3 = K2 * K;
                        b0 = fma_<T>(w, K3, b0);
                        b1 = fma_<T>(wst[i], K3, b1);
                        if constexpr (ADA) b2 = fma_<T>(ws2t[i], K3, b2);
                        if (WG) {
                            const T K5 = K3 * K2;
                            g0 = fma_<T>(w, K5, g0);
                            g1 = fma_<T>(wst[i], K5, g1);
                            g2 = fma_<T>(ws2t[i], K5, g2);
                        }
                    }
                    if (WA) rA = fma_<T>(wct[i], K, rA);
                }
                rBh = b0 - b1;                                  // sum w cos K^3
                if constexpr (ADA) rC2 = b0 - 2 * b1 + b2;      // sum w cos^2 K^3
                rBz = fma_<T>(dm, b0, rf * b1);                 // sum w (rho - r cos) K^3
                rRR = fma_<T>(-dm, g0, fma_<T>(rho + dm, g1, -rho * g2));   // sum w cos (r - rho cos) K^5
                rZZ = fma_<T>(dm, g0, rf * g1);                 // sum w (rho - r cos) K^5
                rRZ = fma_<T>(k0, g0, -fma_<T>(k1, g1, k2 * g2));  // sum w (k0 - k1 s - k2 s^2) K^5
            } else {
#pragma unroll kStableUnroll<T, WA>
                for (int i = 0; i < na; ++i) {
                    const T s = st[i], w = wa[i], wc = wct[i];
                    const T J = fma_<T>(q, s, baseJ);
                    const T K = prsq<P>(J);
                    if (WB || WG) {
                        const T K2 = K * K;
                        const T K3 = K2 * K;
                        const T rmc = fma_<T>(rf, s, dm);     // rho - r cos
                        const T wr = w * rmc;                 // shared by B_z and dB_z/dz
                        rBh = fma_<T>(wc, K3, rBh);        // B_rho, also the gradient's rho-phi term
                        if (WB) rBz = fma_<T>(wr, K3, rBz);
                        if constexpr (ADA) rC2 = fma_<T>(wcst[i], K3, rC2);
                        if (WG) {
                            const T K5 = K3 * K2;
                            // r - rho cos = rho s - dm: two moments, combined per row in double (as in the fast
                            // method; the rho-z term keeps its per-node polynomial, whose moments lose accuracy)
                            sRRs = fma_<T>(wcst[i], K5, sRRs);
                            sRRc = fma_<T>(wc, K5, sRRc);
                            rZZ = fma_<T>(wr, K5, rZZ);
                            const T poly = k0 - s * fma_<T>(k2, s, k1);
                            rRZ = fma_<T>(w, poly * K5, rRZ);
                        }
                    }
                    if (WA) rA = fma_<T>(wc, K, rA);
                }
            }
            const double rowRR = sizeof(T) == 8 ? (double)rRR : fma(rho_d, (double)sRRs, -dd * (double)sRRc);
            const double W = c.cf * wt[j] * rho_d;
            aBh += W * (double)rBh;
            aBz += W * (double)rBz;
            aRR += W * rowRR;
            aZZ += W * (double)rZZ;
            aRZ += W * (double)rRZ;
            aA += W * (double)rA;
            if constexpr (ADA) {   // rows are rings: no length dependence
                const double c2 = sizeof(T) == 8 ? (double)rC2 : (double)rBh - (double)rC2;   // sum w cos^2 K^3
                const double dAdr = -W * (rd * (double)rBh - rho_d * c2);
                const double dAdrho = -(rd * dAdr - zd * zd * W * (double)rBh) / rho_d;
                qR += dAdrho;
                qa += 0.5 * (1.0 + xt[j]) * dAdrho;
                qAr += dAdr;
            }
            if constexpr (AD) {
                const double Wr = c.cf * wt[j], hx = 0.5 * (1.0 + xt[j]);
                const double g0 = zd * ((double)rBh - 3.0 * rd * rowRR + (double)rRZ);           // -d f_rho / d rho
                const double g1 = (double)rBz + rd * (double)rRZ - 3.0 * zd * zd * (double)rZZ;  // -d f_z / d rho
                hR0 += Wr * g0;
                hR1 += Wr * g1;
                ha0 += Wr * hx * g0;
                ha1 += Wr * hx * g1;
            }
        }
        if constexpr (AD) {
            ad->dR[0] = -hR0;
            ad->dR[1] = -hR1;
            ad->da[0] = -ha0;
            ad->da[1] = -ha1;
            ad->db[0] = 0.0;
            ad->db[1] = 0.0;
        }
        if constexpr (ADA) {
            ada->dR = qR;
            ada->da = qa;
            ada->db = 0.0;
            ada->Ar = qAr;
        }
        L.A = aA;
        L.Bh = aBh * zd;
        L.Bz = aBz;
        L.RP = aBh * zd / rd;
        L.RR = -3.0 * zd * aRR;
        L.RZ = aRZ;
        L.ZZ = -3.0 * zd * aZZ;
    }
}

}  // namespace ccoil
