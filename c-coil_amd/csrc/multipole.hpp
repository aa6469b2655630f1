// Far field of a coil as an exterior zonal-harmonic (axial multipole) series, in FP64.
//
// Outside the sphere that holds the coil's winding (radius r_max = |(R + a, b/2)| about its centre), the field of an
// axisymmetric azimuthal current is B = -mu0 grad Phi with Phi = sum_l M_l D^-(l+1) P_l(cos theta), D the distance
// from the centre and theta the angle from the axis. On the axis B_z of a loop of current I, radius rho, height z0
// is mu0 I rho^2 / (2 (rho^2 + (z - z0)^2)^(3/2)); expanding it in c / z (c^2 = rho^2 + z0^2) with the Gegenbauer
// generating function and matching mu0 sum_l (l + 1) M_l z^-(l+2) gives, for current spread over rings (rho, z0)
// with weights w_k summing to N I,
//   M_l = sum_k w_k rho_k^2 V_l(rho_k, z0_k) / (2 (l + 1)),   V_l = c^(l-1) P'_l(z0 / c).
// V_l is a polynomial (V_0 = 0, V_1 = 1, V_(l+1) = c^2 V_(l-1) + (2l + 1) U_l with the solid harmonics
// U_l = c^l P_l(z0 / c), U_(l+1) = ((2l + 1) z0 U_l - l c^2 U_(l-1)) / (l + 1)) of degree at most l - 1 in z0.
//
// The rings are the field rule's own: its thickness nodes (the rule's Gauss-Legendre nodes over [R, R + a], so the
// series reproduces the rule's thickness sum exactly, whatever its order), and over the length either the exact
// average (the fast rule integrates the length in closed form; 126 Gauss-Legendre nodes are exact for l <= 250) or
// the mid-plane alone (the slow rule's flat sheet). What the series leaves out is then the angular rule's own error
// and the truncation. With s = sin theta, u = 1 / D and q_l = (r_max / D)^(l+1), and the moments normalised as
// m_l = mu0 M_l / r_max^(l+1):
//   A_phi = s sum m_l q_l P'_l / l           B_z  = u sum (l + 1) m_l q_l P_(l+1)
//   B_rho = s u sum m_l q_l P'_(l+1)         B_rho / r = u^2 sum m_l q_l P'_(l+1)
//   dB_z/dz = -u^2 sum (l + 1)(l + 2) m_l q_l P_(l+2)
//   dB_rho/dz = dB_z/drho = -s u^2 sum (l + 1) m_l q_l P'_(l+2)
//   dB_rho/drho = -(B_rho / r + dB_z/dz)      (div B = 0)
// The host sets each coil's distance threshold (CoilDev::far_d2) where a bound on the angular rule's error falls
// below the tolerance of the arithmetic that uses the series.
#pragma once

#include "common.hpp"

namespace ccoil {

constexpr int kMpTerms = 250;               // highest order l (below 256, the range of wave_max_small)
constexpr int kMpStride = 256;              // doubles per coefficient array (index l), one per thread of a block
constexpr int kMpCoil = 4 * kMpStride;      // per coil: m_l / l, m_l, (l + 1) m_l, (l + 1)(l + 2) m_l
constexpr int kMpHeader = 3 * kMpStride;    // recurrence coefficients (2l + 5) / (l + 3), (l + 2) / (l + 3), 2l + 5
constexpr int kMpLenNodes = kSeriesLenNodes; // Gauss-Legendre nodes over the length of a fast-rule coil: exact to
                                             // degree 251, and V_l has degree l - 1 in the ring's height
static_assert(kMpTerms < 256 && kMpTerms < kMpStride && kMpStride <= 256, "series orders and their arrays");
static_assert(2 * kMpLenNodes - 1 >= kMpTerms - 1, "the length rule integrates V_l exactly");
// Every coil's winding is symmetric about its mid-plane (the length rule's nodes come in pairs +-z0, a slow-rule
// coil is a mid-plane sheet), so its moments of even order are zero: multipole_moments_kernel writes them as exact
// zeros, and the series loops add only the odd orders.
constexpr bool kMpEvenZero = true;

// Number of terms for a point at distance D from a coil whose winding lies within r_max: the truncation falls like
// (r_max / D)^L with a constant below 2, so L = ln(2 / tol) / ln(D / r_max), between 4 and kMpTerms; log_tol =
// ln(2 / tol): kMpLogTolFP64 (tol = 1e-15, a hundredth of the FP64 rule's own error against the exact integral) or
// kMpLogTolFP32C (1e-13, far below the float rule's own rounding).
constexpr float kMpLogTolFP64 = 35.2319f;
constexpr float kMpLogTolFP32C = 30.6267f;
__device__ inline int multipole_terms(double D2, double inv_rmax2, float log_tol)
{
    const float lg = __logf((float)(D2 * inv_rmax2));   // 2 ln(D / r_max) > 2 ln 1.17
    const int L = (int)(2.0f * log_tol * __builtin_amdgcn_rcpf(lg)) + 2;
    return L < 4 ? 4 : (L > kMpTerms ? kMpTerms : L);
}

// Largest v over the wave for 0 <= v < 256 (every lane active): its bits from the top, one ballot each.
__device__ inline int wave_max_small(int v)
{
    int m = 0;
#pragma unroll
    for (int b = 128; b >= 1; b >>= 1)
        if (__ballot(v >= (m | b))) m |= b;
    return m;
}

// Smallest v over the lanes with active set, for 0 <= v < 256 (every lane of the wave executing); 255 if none is.
__device__ inline int wave_min_small(int v, bool active)
{
    return 255 - wave_max_small(active ? 255 - v : 0);
}

// Local (A_phi, B_rho, B_z, B_rho / r, dB_rho/drho, dB_rho/dz, dB_z/dz) of one coil at local (z, r) from its series
// (mp: the coil's kMpCoil doubles, rec: the shared recurrence coefficients, rmax: the coil's r_max): the loop runs L
// terms (wave-uniform), and terms beyond own add exact zeros, so the value is that of own terms. The first lo terms
// (the least own count among the lanes that call; wave-uniform) need no such test.
template <bool WA, bool WB, bool WG>
__device__ inline void eval_multipole(const double *mp, const double *rec, double rmax, double z, double r, int L,
                                      int lo, int own, Local &out)
{
    const double D2 = r * r + z * z;
    const double u = rsqrt_newton(D2);
    const double c = z * u, s = r * u, qr = rmax * u;
    const double *mA = mp, *mB = mp + kMpStride, *mZ = mp + 2 * kMpStride, *mZZ = mp + 3 * kMpStride;
    // P_(l+1), P_(l+2) and P'_l, P'_(l+1), P'_(l+2), advanced together from l = 1
    double P1 = 0.5 * (3.0 * c * c - 1.0), P2 = 0.5 * c * (5.0 * c * c - 3.0);
    double d0 = 1.0, d1 = 3.0 * c, d2 = 1.5 * (5.0 * c * c - 1.0);
    double q = qr;   // q_l = (r_max / D)^(l + 1)
    double SA = 0.0, SZ = 0.0, SB = 0.0, SZZ = 0.0, SRZ = 0.0;
    auto accumulate = [&](int l, double ql) {
        if (WA) SA = fma(mA[l] * ql, d0, SA);
        if (WB || WG) {
            SZ = fma(mZ[l] * ql, P1, SZ);
            SB = fma(mB[l] * ql, d1, SB);
        }
        if (WG) {
            SZZ = fma(mZZ[l] * ql, P2, SZZ);
            SRZ = fma(mZ[l] * ql, d2, SRZ);
        }
    };
    auto step = [&](int l) {
        // P_(l+3) = ((2l + 5) c P_(l+2) - (l + 2) P_(l+1)) / (l + 3), P'_(l+3) = P'_(l+1) + (2l + 5) P_(l+2)
        const double P3 = fma(rec[l] * c, P2, -rec[kMpStride + l] * P1);
        const double d3 = fma(rec[2 * kMpStride + l], P2, d1);
        P1 = P2;
        P2 = P3;
        d0 = d1;
        d1 = d2;
        d2 = d3;
    };
    // The moments of even order are exactly zero (kMpEvenZero), so only odd l accumulate while the recurrences and
    // the powers q_l step through every l: the same operations on the same values as summing every term.
    int l = 1;
#pragma unroll 2
    for (; l <= lo; l += 2) {
        q *= qr;
        accumulate(l, q);
        step(l);
        q *= qr;
        step(l + 1);
    }
#pragma unroll 2
    for (; l <= L; l += 2) {
        q *= qr;
        accumulate(l, l <= own ? q : 0.0);
        step(l);
        q *= qr;
        step(l + 1);
    }
    out.A = s * SA;
    out.Bz = u * SZ;
    out.Bh = s * u * SB;
    out.RP = u * u * SB;
    out.ZZ = -u * u * SZZ;
    out.RZ = -s * u * u * SRZ;
    out.RR = -(out.RP + out.ZZ);
}

// The pair kernel's view of the series at a secondary node: A_phi (WA), B_rho and B_z (WB), and A_phi / r = u sum
// m_l q_l P'_l / l (also on the axis, where it is the limit), for dA_phi/dr = B_z - A_phi / r. Loop as in
// eval_multipole: L terms, the first lo without the test, the lane's own count own.
template <bool WA, bool WB>
__device__ inline void eval_multipole_pair(const double *mp, const double *rec, double rmax, double z, double r, int L,
                                           int lo, int own, Local &out, double &a_over_r)
{
    const double D2 = r * r + z * z;
    const double u = rsqrt_newton(D2);
    const double c = z * u, s = r * u, qr = rmax * u;
    const double *mA = mp, *mB = mp + kMpStride, *mZ = mp + 2 * kMpStride;
    double P1 = 0.5 * (3.0 * c * c - 1.0), P2 = 0.5 * c * (5.0 * c * c - 3.0);
    double d0 = 1.0, d1 = 3.0 * c, d2 = 1.5 * (5.0 * c * c - 1.0);
    double q = qr;
    double SA = 0.0, SZ = 0.0, SB = 0.0;
    auto accumulate = [&](int l, double ql) {
        if (WA) SA = fma(mA[l] * ql, d0, SA);
        if (WB) {
            SZ = fma(mZ[l] * ql, P1, SZ);
            SB = fma(mB[l] * ql, d1, SB);
        }
    };
    auto step = [&](int l) {
        const double P3 = fma(rec[l] * c, P2, -rec[kMpStride + l] * P1);
        const double d3 = fma(rec[2 * kMpStride + l], P2, d1);
        P1 = P2;
        P2 = P3;
        d0 = d1;
        d1 = d2;
        d2 = d3;
    };
    int l = 1;   // odd l only, as in eval_multipole
#pragma unroll 2
    for (; l <= lo; l += 2) {
        q *= qr;
        accumulate(l, q);
        step(l);
        q *= qr;
        step(l + 1);
    }
#pragma unroll 2
    for (; l <= L; l += 2) {
        q *= qr;
        accumulate(l, l <= own ? q : 0.0);
        step(l);
        q *= qr;
        step(l + 1);
    }
    out.A = s * SA;
    out.Bz = u * SZ;
    out.Bh = s * u * SB;
    a_over_r = u * SA;
}

}  // namespace ccoil
