// Adjoint of the summed B field with respect to every coil's parameters (gfx950).
//
// For a scalar L of the summed field, the caller passes g(x) = dL/dB(x); this computes, per coil c,
//   dL/dtheta_c = sum_x g(x) . dB_c(x)/dtheta_c   for theta = (I, R, a, b, position, alpha, beta).
// With B_c(x) = T B_loc(T^T (x - p)) (local frame, CalculateFieldsMethods.cxx:16-19):
//   d/dI        B_c / I                              (B is linear in the current through J = N I / (a b))
//   d/dR, a, b  T dB_loc/d(R, a, b)                  (homogeneity of the quadrature rows, AdjLocal)
//   d/dp        -J^T g with J = T G_loc T^T          (translation)
//   d/dtheta    g_loc . (Omega B_loc - G_loc Omega x_loc), Omega = T^T dT/dtheta, g_loc = T^T g
// The quadrature rule (PrecisionArguments) is held fixed, as the reference's integer planning is piecewise
// constant in the geometry. One thread per (coil, point); coils are wave-uniform (grid y), partial sums per
// (coil, block) are reduced in a fixed order, so results are deterministic.
#include "integrands.hpp"
#include "kernels.hpp"

namespace ccoil {

constexpr int kGradBlock = 256;

__device__ inline double wave_reduce(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ inline void cross(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

template <int PREC>
__global__ __launch_bounds__(kGradBlock) void field_grad_kernel(GradParams p)
{
    const int ci = blockIdx.y;
    const CoilDev &c = p.coils[ci];
    double acc[kGradTerms];
    for (int k = 0; k < kGradTerms; ++k) acc[k] = 0.0;

    for (int64_t pi = (int64_t)blockIdx.x * kGradBlock + threadIdx.x; pi < p.n_pts;
         pi += (int64_t)p.n_blocks * kGradBlock) {
        double z, r, cph, sph;
        to_local(c, p.pts[3 * pi], p.pts[3 * pi + 1], p.pts[3 * pi + 2], z, r, cph, sph);
        Local L = {0, 0, 0, 0, 0, 0, 0};
        AdjLocal ad;
        if constexpr (PREC == kFP64) {
            eval_stable<f64s, false, true, true, true>(c, p.tab, z, r, 0, 1, L, &ad);
        } else if constexpr (PREC == kFP64IEEE) {
            eval_stable<double, false, true, true, true>(c, p.tab, z, r, 0, 1, L, &ad);
        } else {
            // a wave holding a near-winding or near-axis pair evaluates all its lanes in FP64: a float pass beside
            // the FP64 one would add its own issue time. A pair's value thus depends on its wave, which the
            // point order and the grid fix, as they fix the order of the sums over points.
            if (__ballot(is_near(c, z, r, true)))
                eval_stable<f64s, false, true, true, true>(c, p.tab, z, r, 0, 1, L, &ad);
            else
                eval_stable<float, false, true, true, true>(c, p.tab, z, r, 0, 1, L, &ad);
        }
        // local Cartesian B, gradient (the same on-axis rule as the forward path) and position
        const double Bl[3] = {L.Bh * cph, L.Bh * sph, L.Bz};
        double G[9];
        if (r <= kGpuAxisRatio * c.R) {
            G[0] = L.RR; G[1] = 0.0; G[2] = 0.0;
            G[3] = 0.0; G[4] = L.RR; G[5] = 0.0;
            G[6] = 0.0; G[7] = 0.0; G[8] = L.ZZ;
        } else {
            const double xy = cph * sph * (L.RR - L.RP);
            G[0] = L.RR * cph * cph + L.RP * sph * sph; G[1] = xy; G[2] = L.RZ * cph;
            G[3] = xy; G[4] = L.RR * sph * sph + L.RP * cph * cph; G[5] = L.RZ * sph;
            G[6] = L.RZ * cph; G[7] = L.RZ * sph; G[8] = L.ZZ;
        }
        const double xl[3] = {r * cph, r * sph, z};
        const double *gw = p.gB + 3 * pi;
        double gl[3];
        mat_vec(c.Tinv, gw, gl);
        const double grho = gl[0] * cph + gl[1] * sph, gz = gl[2];
        // dI: B is proportional to the current the constFactor was formed with
        acc[0] += (gl[0] * Bl[0] + gl[1] * Bl[1] + gl[2] * Bl[2]) / c.cur_used;
        acc[1] += grho * ad.dR[0] + gz * ad.dR[1];
        acc[2] += grho * ad.da[0] + gz * ad.da[1];
        acc[3] += grho * ad.db[0] + gz * ad.db[1];
        // position: -T (G_loc g_loc) (G_loc symmetric off the axis; on it the reference's diagonal form)
        double Gg[3];
        mat_vec(G, gl, Gg);
        double dp[3];
        mat_vec(c.T, Gg, dp);
        acc[4] -= dp[0];
        acc[5] -= dp[1];
        acc[6] -= dp[2];
        const double *om[2] = {c.om_a, c.om_b};
        for (int k = 0; k < 2; ++k) {
            double oB[3], ox[3], Gox[3];
            cross(om[k], Bl, oB);
            cross(om[k], xl, ox);
            mat_vec(G, ox, Gox);
            acc[7 + k] += gl[0] * (oB[0] - Gox[0]) + gl[1] * (oB[1] - Gox[1]) + gl[2] * (oB[2] - Gox[2]);
        }
    }

    __shared__ double red[kGradBlock / 64][kGradTerms];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int k = 0; k < kGradTerms; ++k) {
        const double v = wave_reduce(acc[k]);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kGradTerms) {
        const int k = threadIdx.x;
        double s = 0.0;
        for (int w = 0; w < kGradBlock / 64; ++w) s += red[w][k];
        if (k > 0) s *= c.adj_scale;
        p.partial[((int64_t)ci * p.n_blocks + blockIdx.x) * kGradTerms + k] = s;
    }
}

// grad[c][k] = sum over blocks, in block order.
__global__ void field_grad_reduce_kernel(GradParams p)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)p.n_coils * kGradTerms) return;
    const int64_t ci = t / kGradTerms, k = t % kGradTerms;
    double s = 0.0;
    for (int b = 0; b < p.n_blocks; ++b) s += p.partial[(ci * p.n_blocks + b) * kGradTerms + k];
    p.grad[t] = s;
}

void launch_field_grad(hipStream_t s, const GradParams &p, int precision)
{
    if (p.n_coils <= 0) return;
    const dim3 grid((unsigned)p.n_blocks, (unsigned)p.n_coils);
    if (p.n_pts > 0) {
        if (precision == kFP64) hipLaunchKernelGGL(field_grad_kernel<kFP64>, grid, dim3(kGradBlock), 0, s, p);
        else if (precision == kFP64IEEE)
            hipLaunchKernelGGL(field_grad_kernel<kFP64IEEE>, grid, dim3(kGradBlock), 0, s, p);
        else hipLaunchKernelGGL(field_grad_kernel<kFP32C>, grid, dim3(kGradBlock), 0, s, p);
    }
    const int64_t n = (int64_t)p.n_coils * kGradTerms;
    hipLaunchKernelGGL(field_grad_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p);
}

}  // namespace ccoil
