// Coaxial (z-axis) mutual inductance, axial force and self-inductance (gfx950).
//
// The length integrals of both coils are analytic in these rules; what remains is a 3-D Gauss-Legendre sum over
// the secondary's radial nodes r, the primary's thickness nodes t and the primary's angular nodes phi:
//   mutual  CalculateMInductanceZAxis.cxx:84-221   sum of w * 2tr cos(phi) * (H1 - H2 + H3 - H4 - E1 + E2 - E3 + E4)
//   force   CalculateForceZAxis.cxx:82-219         sum of w * 2tr cos(phi) * log(G2 G4 / (G1 G3))
//   self    CalculateMInductanceZAxis.cxx:224-317  sum of w * 2tr cos(phi) * (D - E + l H)
// One thread owns one (r, phi) unit and loops over the t nodes; the per-node terms are evaluated exactly as the
// reference writes them, in double with IEEE division and square root and without contraction, so the result
// differs from the reference's CPU sum only by the summation order and the last bit of cos and log.
#include "zaxis.hpp"

namespace ccoil {

namespace {

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

__global__ __launch_bounds__(kZBlock) void zaxis_kernel(ZParams p)
{
#pragma clang fp contract(off)
    __shared__ double wsum[kZBlock / 64];
    const int64_t wg = blockIdx.x;
    // the item whose workgroup range holds wg: last item with first_wg <= wg (workgroup-uniform search)
    int lo = 0, hi = p.n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.items[mid].first_wg <= wg) lo = mid;
        else hi = mid - 1;
    }
    const ZItemDev &it = p.items[lo];
    const int64_t u = (wg - it.first_wg) * kZBlock + threadIdx.x;
    double acc = 0.0;
    if (u < it.units) {
        const int n_phi = it.a_blk * it.a_inc;
        const int ip = (int)(u % n_phi), ir = (int)(u / n_phi);
        const int ab = ip / it.a_inc, ai = ip - ab * it.a_inc;
        const int rb = ir / it.r_inc, ri = ir - rb * it.r_inc;
        const double *xa = p.tab.x + table_offset(it.a_inc), *wa = p.tab.w + table_offset(it.a_inc);
        const double *xr = p.tab.x + table_offset(it.r_inc), *wr = p.tab.w + table_offset(it.r_inc);
        const double *xt = p.tab.x + table_offset(it.t_inc), *wt = p.tab.w + table_offset(it.t_inc);
        const double cp = cos(it.a_width * (ab + 0.5) + (it.a_width * 0.5) * xa[ai]);
        const double wF = wa[ai];
        const double r = (it.r_inner + it.r_width * (rb + 0.5)) + (it.r_width * 0.5) * xr[ri];
        const double wR = wr[ri];
        for (int tb = 0; tb < it.t_blk; ++tb) {
            const double tpos = it.t_inner + it.t_width * (tb + 0.5);
            for (int ti = 0; ti < it.t_inc; ++ti) {
                const double t = tpos + (it.t_width * 0.5) * xt[ti];
                const double tc = it.constant * wR * wt[ti];
                const double A = 2.0 * t * r;
                const double B = t * t + r * r;
                const double C = B - A * cp;
                const double D = sqrt(C);
                double f;
                if (it.mode == kZMutual) {
                    double E[4], H[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        E[k] = sqrt(C + it.zsq[k]);
                        const double F = it.z[k] / D;
                        const double G = sqrt(F * F + 1.0);
                        H[k] = it.z[k] * log(F + G);
                    }
                    f = H[0] - H[1] + H[2] - H[3] - E[0] + E[1] - E[2] + E[3];
                } else if (it.mode == kZForce) {
                    double G[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const double E = it.z[k] / D;
                        G[k] = E + sqrt(E * E + 1.0);
                    }
                    f = log(G[1] * G[3] / (G[0] * G[2]));
                } else {
                    const double len = it.z[0];
                    const double E = sqrt(C + it.zsq[0]);
                    const double F = len / D;
                    const double G = sqrt(F * F + 1.0);
                    f = D - E + len * log(F + G);
                }
                acc += tc * wF * A * cp * f;
            }
        }
    }
    // fixed-order workgroup sum: butterfly within each wave, then the waves in index order
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kZBlock / 64; ++w) s += wsum[w];
        p.partial[wg] = s;
    }
}

// One thread per item: the item's workgroup partials in workgroup order, then the reference's final factors.
__global__ __launch_bounds__(256) void zaxis_reduce_kernel(ZParams p)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= p.n_items) return;
    const ZItemDev &it = p.items[k];
    const int64_t end = (k + 1 < p.n_items) ? p.items[k + 1].first_wg : p.n_wg;
    double s = 0.0;
    for (int64_t w = it.first_wg; w < end; ++w) s += p.partial[w];
    p.out[it.out] = ((s * it.mul) * it.mul2) / it.div;
}

void launch_zaxis(hipStream_t s, const ZParams &p)
{
    if (p.n_wg <= 0) return;
    hipLaunchKernelGGL(zaxis_kernel, dim3((unsigned)p.n_wg), dim3(kZBlock), 0, s, p);
}

void launch_zaxis_reduce(hipStream_t s, const ZParams &p)
{
    if (p.n_items <= 0) return;
    hipLaunchKernelGGL(zaxis_reduce_kernel, dim3((p.n_items + 255) / 256), dim3(256), 0, s, p);
}

}  // namespace ccoil
