// Device-side data layout shared by the field, pair and table kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccoil {

constexpr int kMaxOrder = 100;                        // Legendre::maxLegendreOrder (LegendreMatrix.h:9)
constexpr int kTableSize = kMaxOrder * (kMaxOrder + 1) / 2;
// The far-field series' length nodes: a Gauss-Legendre rule of kSeriesLenNodes nodes (exact to degree 251), stored in
// the x and w tables after the orders 1..kMaxOrder, at kSeriesLenOffset.
constexpr int kSeriesLenNodes = 126;
constexpr int kSeriesLenOffset = kTableSize;
constexpr int kTableSizeXW = kTableSize + kSeriesLenNodes;
constexpr double kPi = 3.14159265358979323846;
constexpr double kMuReduced = 1e-7;                    // mu0 / (4 pi), g_MiReduced (CalculateFieldsFast.cxx:10)
constexpr double kGpuAxisRatio = 1e-5;                 // on-axis branch of the GPU gradient (FieldGGroup.cu:134)

__host__ __device__ inline int table_offset(int n) { return n * (n - 1) / 2; }

// One coil as the kernels see it (FP64). Indexed with a wave-uniform coil index, so every field access
// is a scalar (SMEM) load broadcast to the wave.
struct alignas(64) CoilDev {
    double pos[3];
    double T[9];      // local -> global, row-major
    double Tinv[9];   // global -> local
    double R, a, b;   // inner radius, thickness, length (clamped)
    double cf;        // constFactor of generateCoilData (SupportFunctions.cxx:169-172)
    double omega;     // 2 pi f (E = omega * A)
    double current;   // I
    double turns;     // N
    double near_w2;   // squared near-winding shell width
    double sa, ca, sb, cb;   // sin/cos of alpha (y-axis angle) and beta (z-axis angle)
    double om_a[3], om_b[3]; // T^T dT/dalpha and T^T dT/dbeta as axial vectors (Omega v = om x v)
    double cur_used;         // current the constFactor was formed with (1 for a zero-current coil)
    double adj_scale;        // I / cur_used: scales every adjoint term except d/dI
    double ni;               // N I: the coil's total current (J a b)
    double rmax;             // radius of the sphere about the centre that holds the winding, |(R + a, b / 2)|
    double far_d2;           // squared distance beyond which the far-field series replaces the rule (0: never)
    double inv_rmax2;        // 1 / rmax^2
    int32_t use_fast;
    int32_t n_ang, n_thk, n_len;
    int32_t blk_ang, blk_thk, blk_len;
    int32_t near_key; // FP32C near pass: the coil's place in the evaluation order (most nodes first)
    int32_t mp_terms; // highest order of the far-field series a pair of this call can use (0: no series)
};

// Gauss-Legendre tables for orders 1..100 at offset table_offset(n), generated in FP64 on the device (x and w also
// hold the series' length rule at kSeriesLenOffset).
struct Tables {
    const double *x;      // nodes, ascending
    const double *w;      // weights
    const double *cosv;   // cos(pi/2 (1 + x))        (SupportFunctions.cxx:188)
    const double *omc;    // 1 - cos(pi/2 (1 + x)) = 2 sin^2(pi/4 (1 + x))
    const float *xf;
    const float *wf;
    const float *cosf;
    const float *omcf;
    const double *rcos;   // cos(pi (1 + x)): secondary ring nodes of the GPU pair rule (MInductance.cu:34)
    const double *rsin;   // sin(pi (1 + x))
    const double *wcos;   // w cos(pi/2 (1 + x)): the angular weight times the cosine
    const float *wcosf;
    const double *opc;    // 1 + cos(pi/2 (1 + x)) = 2 cos^2(pi/4 (1 + x))
    const float *opcf;
    const double *ws;     // w (1 - cos): the weight times the first and second power of s = 1 - cos
    const double *ws2;    // w (1 - cos)^2
    const double *wcs;    // w cos (1 - cos)
    const float *wcsf;
};

// Quantity bits (mirror of the C ABI enum).
enum : uint32_t { Q_A = 1, Q_B = 2, Q_E = 4, Q_G = 8, Q_GJ = 16 };

// Local cylindrical results of one (coil, point): A_phi, B_rho, B_z and the four gradient terms of
// CalculateFieldsMethods.cxx:53-56 (RPhi, RR, RZ, ZZ).
struct Local {
    double A, Bh, Bz, RP, RR, RZ, ZZ;
};

struct Cart {
    double A[3], E[3], B[3], G[9];
};

__device__ inline void mat_vec(const double *M, const double *v, double *o)
{
    o[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
    o[1] = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
    o[2] = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
}

// 1 / sqrt(x) for a finite positive normal x: v_rsq_f64's estimate (about 2^-23 relative) refined by two Newton steps
// on f(y) = 1 / y^2 - x, y <- y + (y / 2) (1 - x y^2), each of which squares the relative error; within an ulp or two
// of the exact value and a fixed sequence of operations, so the same in every kernel.
__device__ inline double rsqrt_newton(double x)
{
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-(x * y), y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// Point -> coil-local cylindrical (z, r) and the azimuth's cos/sin (cos(atan2(y,x)) = x/r).
// Every operation is rounded on its own (no FMA contraction), so each kernel that classifies a (coil, point) pair
// with is_near gets bitwise the same z and r from the same inputs, whatever the inlining context.
__device__ inline void to_local(const CoilDev &c, double px, double py, double pz, double &z, double &r,
                                double &cph, double &sph)
{
#pragma clang fp contract(off)
    const double d[3] = {px - c.pos[0], py - c.pos[1], pz - c.pos[2]};
    const double *M = c.Tinv;   // written out here rather than through mat_vec, so the pragma covers it
    const double l[3] = {M[0] * d[0] + M[1] * d[1] + M[2] * d[2], M[3] * d[0] + M[4] * d[1] + M[5] * d[2],
                         M[6] * d[0] + M[7] * d[1] + M[8] * d[2]};
    z = l[2];
    const double r2 = l[0] * l[0] + l[1] * l[1];
    if (r2 >= 1e-200) {
        // r and 1 / r from one reciprocal square root (within about an ulp of the correctly rounded values)
        const double ir = rsqrt_newton(r2);
        r = r2 * ir;
        cph = l[0] * ir;
        sph = l[1] * ir;
    } else if (r2 > 0.0) {   // within 1e-100 of the axis: the IEEE operations, clear of the estimates' range limits
        r = sqrt(r2);
        cph = l[0] / r;
        sph = l[1] / r;
    } else {
        r = 0.0;
        cph = 1.0;
        sph = 0.0;
    }
}

// Local cylindrical -> global Cartesian, accumulated into out (CalculateFieldsMethods.cxx:16-19,
// :33-36, :58-83; FieldGGroup.cu:134-174 for the GPU on-axis threshold).
__device__ inline void accumulate_global(const CoilDev &c, const Local &L, double r, double cph, double sph,
                                         uint32_t q, Cart &out)
{
    if (q & (Q_A | Q_E)) {
        double v[3] = {-L.A * sph, L.A * cph, 0.0}, g[3];
        mat_vec(c.T, v, g);
        if (q & Q_A)
            for (int k = 0; k < 3; ++k) out.A[k] += g[k];
        if (q & Q_E)
            for (int k = 0; k < 3; ++k) out.E[k] += c.omega * g[k];
    }
    if (q & Q_B) {
        double v[3] = {L.Bh * cph, L.Bh * sph, L.Bz}, g[3];
        mat_vec(c.T, v, g);
        for (int k = 0; k < 3; ++k) out.B[k] += g[k];
    }
    if (q & (Q_G | Q_GJ)) {
        double g[9];
        if (r <= kGpuAxisRatio * c.R) {   // r / R <= 1e-5 without the division
            g[0] = L.RR; g[1] = 0.0; g[2] = 0.0;
            g[3] = 0.0; g[4] = L.RR; g[5] = 0.0;
            g[6] = 0.0; g[7] = 0.0; g[8] = L.ZZ;
        } else {
            double xx = L.RR * cph * cph + L.RP * sph * sph;
            double yy = L.RR * sph * sph + L.RP * cph * cph;
            double xy = cph * sph * (L.RR - L.RP);   // 0.5 sin(2 phi) (RR - RP)
            double xz = L.RZ * cph, yz = L.RZ * sph;
            g[0] = xx; g[1] = xy; g[2] = xz;
            g[3] = xy; g[4] = yy; g[5] = yz;
            g[6] = xz; g[7] = yz; g[8] = L.ZZ;
        }
        double tg[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                tg[3 * i + j] = c.T[3 * i] * g[j] + c.T[3 * i + 1] * g[3 + j] + c.T[3 * i + 2] * g[6 + j];
        if (q & Q_GJ) {
            // (T G) T^T
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    out.G[3 * i + j] += tg[3 * i] * c.T[3 * j] + tg[3 * i + 1] * c.T[3 * j + 1] +
                                        tg[3 * i + 2] * c.T[3 * j + 2];
        } else {
            for (int k = 0; k < 9; ++k) out.G[k] += tg[k];
        }
    }
}

// Pairs that FP32C evaluates in FP64:
//  - the near-winding shell: distance from (r, z) to the cross-section rectangle [R, R+a] x [-b/2, b/2] below one
//    cross-section width (max(a, b), floored at 1e-3 R for filaments);
//  - with axis (outputs that divide by r: the gradient's hoop term B_rho / r, the pair gradient's A / r), the cone
//    and cylinder around the axis, r < 0.1 sqrt(r^2 + z^2 + R^2): the float angular sums of B_rho and A vanish like
//    r while their rounding does not.
constexpr double kAxisCone2 = 0.01;

// Whether the far-field series replaces the quadrature rule for a point at local (z, r). Rounded operation by
// operation, like to_local, so every kernel that classifies a pair gets the same answer.
__device__ inline bool is_far(const CoilDev &c, double z, double r)
{
#pragma clang fp contract(off)
    return c.far_d2 > 0.0 && r * r + z * z > c.far_d2;
}

__device__ inline bool is_near(const CoilDev &c, double z, double r, bool axis)
{
#pragma clang fp contract(off)
    double dr = fmax(fmax(c.R - r, r - (c.R + c.a)), 0.0);
    double dz = fmax(fabs(z) - 0.5 * c.b, 0.0);
    const double r2 = r * r;
    return dr * dr + dz * dz < c.near_w2 || (axis && r2 < kAxisCone2 * (r2 + z * z + c.R * c.R));
}

// The pairs of the FP32C near pass: near and not far (the far-field series is FP64 and exact on the axis).
__device__ inline bool near_pass_pair(const CoilDev &c, double z, double r, bool axis)
{
    return !is_far(c, z, r) && is_near(c, z, r, axis);
}

}  // namespace ccoil
