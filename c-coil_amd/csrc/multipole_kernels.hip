// Far-field moments of the coils and the spatial visit order of the points (gfx950).
//
// multipole_moments_kernel: one workgroup per coil sums the zonal-harmonic moments of the coil's rings (the field
// rule's thickness nodes times the length nodes, multipole.hpp) in FP64, with a fixed reduction order.
//
// The visit order: the points' bounding box, a cell of a 2^k x 2^k x 2^k grid for every point (Morton order of the
// cells), a counting sort of the point indices by cell. field_kernel reads its points in that order, so the 64
// points of a wave lie in one or two neighbouring cells and a wave mostly takes one branch (series or rule) per
// coil. Each point's arithmetic does not depend on which wave holds it, so the order within a cell (set by
// atomics) does not change any result.
#include "kernels.hpp"
#include "multipole.hpp"

namespace ccoil {

constexpr int kMpChunks = kMpStride / 64;   // blocks of 64 orders per coil
constexpr int kMomThreads = 512;            // threads of a moments block (8 waves; at most 2 per SIMD by registers)
constexpr int kMomWaves = kMomThreads / 64;
static_assert(kMpChunks == 4, "multipole_moments_kernel dispatches four chunks");

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Adds w V_l of one ring (radius^2 c2, height z0, both in units of r_max) to acc[l - L0] for the orders l of the
// chunk [L0, L0 + 64) up to kMpTerms. The recurrence runs from l = 1 with compile-time indices, so acc stays in
// registers (each chunk repeats the recurrence below it).
template <int L0>
__device__ inline void ring_moments(double w, double z0, double c2, double *acc)
{
    constexpr int kEnd = L0 + 64 < kMpTerms + 1 ? L0 + 64 : kMpTerms + 1;   // orders below kEnd
    double Um = 1.0, U = z0;     // U_(l-1), U_l
    double Vm = 0.0, V = 1.0;    // V_(l-1), V_l
    if constexpr (L0 <= 1) acc[1 - L0] += w;
#pragma unroll
    for (int l = 1; l + 1 < kEnd; ++l) {
        const double Un = ((2.0 * l + 1.0) * z0 * U - l * c2 * Um) * (1.0 / (l + 1.0));
        const double Vn = c2 * Vm + (2.0 * l + 1.0) * U;
        Um = U;
        U = Un;
        Vm = V;
        V = Vn;
        if (l + 1 >= L0) acc[l + 1 - L0] += w * V;
    }
}

// The moments of orders [L0, L0 + 64) of coil c into out (a block of kMomThreads threads over the coil's rings).
// The length rule is symmetric about the mid-plane and V_l(rho, -z0) = (-1)^(l-1) V_l(rho, z0) (every operation of
// the recurrence flips sign exactly with z0), so a node pair adds 2 V_l for odd l and exactly nothing for even l:
// the rings take the nodes z0 > 0 at twice their weight, and the even orders of a coil with a length rule are zero.
// (A slow-rule coil's mid-plane ring has V_l = 0 for even l by itself.)
template <int L0>
__device__ inline void chunk_moments(const CoilDev &c, const Tables &tab, double *out, double (*part)[64])
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const double inv = 1.0 / c.rmax;
    static_assert(kMpLenNodes % 2 == 0, "the length rule's nodes come in pairs +-z0");
    const int nt = c.n_thk, nz = c.use_fast ? kMpLenNodes / 2 : 1;
    const double *xt = tab.x + table_offset(nt), *wt = tab.w + table_offset(nt);
    // the upper half of the ascending length nodes: z0 > 0
    const double *xz = tab.x + kSeriesLenOffset + kMpLenNodes / 2, *wz = tab.w + kSeriesLenOffset + kMpLenNodes / 2;
    double acc[64];
#pragma unroll
    for (int l = 0; l < 64; ++l) acc[l] = 0.0;
    for (int k = t; k < nt * nz; k += kMomThreads) {
        const int j = k / nz, i = k - j * nz;
        // ring (rho, z0) in units of r_max, weight: its share of the current (the weights sum to 1)
        const double rho = (c.R + 0.5 * c.a * (1.0 + xt[j])) * inv;
        const double z0 = nz > 1 ? 0.5 * c.b * xz[i] * inv : 0.0;
        const double w = 0.5 * wt[j] * (nz > 1 ? wz[i] : 1.0) * rho * rho;   // 2 x (1/2) wz for the node pair
        ring_moments<L0>(w, z0, rho * rho + z0 * z0, acc);
    }
#pragma unroll
    for (int l = 0; l < 64; ++l) {
        const double v = wave_sum(acc[l]);
        if (lane == 0) part[wave][l] = v;
    }
    __syncthreads();
    // m_l = mu0 N I sum / (2 (l + 1)), mu0 = 4 pi 1e-7
    const double ni = c.ni * (4.0 * kPi * kMuReduced);
    if (t < 64 && L0 + t < kMpStride) {
        const int l = L0 + t;
        double m = 0.0;
        if (l >= 1 && l <= kMpTerms && (l % 2 == 1 || !c.use_fast)) {
            const double s = ((part[0][t] + part[1][t]) + (part[2][t] + part[3][t])) +
                             ((part[4][t] + part[5][t]) + (part[6][t] + part[7][t]));
            m = ni * s / (2.0 * (l + 1));
        }
        out[l] = l >= 1 ? m / l : 0.0;
        out[kMpStride + l] = m;
        out[2 * kMpStride + l] = (l + 1) * m;
        out[3 * kMpStride + l] = (l + 1.0) * (l + 2.0) * m;
    }
    __syncthreads();
}

// Grid (coils, kMpChunks): block (c, k) sums the orders [64 k, 64 k + 64) of coil c, or zeroes them when the call's
// series of that coil stops below them (mp_terms).
__global__ __launch_bounds__(kMomThreads) void multipole_moments_kernel(const CoilDev *coils, Tables tab, double *mp)
{
    static_assert(kMomWaves == 8, "chunk_moments sums eight wave partials");
    __shared__ double part[kMomWaves][64];
    const int ci = blockIdx.x, chunk = blockIdx.y, t = threadIdx.x;
    if (ci == 0 && chunk == 0 && t < kMpStride) {   // recurrence coefficients of the Legendre polynomials
        mp[t] = (2.0 * t + 5.0) / (t + 3.0);
        mp[kMpStride + t] = (t + 2.0) / (t + 3.0);
        mp[2 * kMpStride + t] = 2.0 * t + 5.0;
    }
    const CoilDev &c = coils[ci];
    double *out = mp + kMpHeader + (size_t)ci * kMpCoil;
    const int l0 = 64 * chunk;
    if (!(c.far_d2 > 0.0) || (chunk > 0 && c.mp_terms < l0)) {   // no series, or none of these orders
        const int l = l0 + t;
        if (t < 64 && l < kMpStride)
            for (int k = 0; k < 4; ++k) out[k * kMpStride + l] = 0.0;
        return;
    }
    if (chunk == 0) chunk_moments<0>(c, tab, out, part);
    else if (chunk == 1) chunk_moments<64>(c, tab, out, part);
    else if (chunk == 2) chunk_moments<128>(c, tab, out, part);
    else chunk_moments<192>(c, tab, out, part);
}

// ---------------------------------------------------------------------------------------------
// Visit order.
// ---------------------------------------------------------------------------------------------
constexpr int kBoxBlocks = 512;

__device__ inline void minmax6(double *v, const double *o)
{
    for (int k = 0; k < 3; ++k) {
        v[k] = fmin(v[k], o[k]);
        v[3 + k] = fmax(v[3 + k], o[3 + k]);
    }
}

// Block-wide min/max of 6 values (3 minima, 3 maxima), result in thread 0.
__device__ inline void block_minmax6(double *v)
{
    __shared__ double sh[4][6];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        double o[6];
        for (int k = 0; k < 6; ++k) o[k] = __shfl_xor(v[k], m, 64);
        minmax6(v, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 6; ++k) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < 4; ++w) minmax6(v, sh[w]);
}

__global__ __launch_bounds__(256) void box_partial_kernel(const double *pts, int64_t n, double *part)
{
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double p[6] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        minmax6(v, p);
    }
    block_minmax6(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = v[k];
}

__global__ __launch_bounds__(256) void box_final_kernel(const double *part, int n_part, double *box)
{
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < n_part; b += 256) minmax6(v, part + 6 * b);
    block_minmax6(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) box[k] = v[k];
}

// Spreads the low 10 bits of x to every third bit.
__device__ inline uint32_t spread3(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Cell (Morton code of the 2^bits grid over the box) of every point, and the points per cell.
__global__ __launch_bounds__(256) void cell_count_kernel(const double *pts, int64_t n, const double *box, int bits,
                                                         uint32_t *cell, unsigned long long *count)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int cells = 1 << bits;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const double ext = box[3 + k] - box[k];
        const double f = ext > 0.0 ? (pts[3 * i + k] - box[k]) / ext * cells : 0.0;   // in [0, cells]
        q[k] = f >= 1.0 ? (uint32_t)fmin(f, cells - 1.0) : 0u;   // NaN and non-finite boxes: cell 0
    }
    const uint32_t key = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    cell[i] = key;
    atomicAdd(&count[key], 1ull);
}

// order[cursor[cell]++] = point (cursor: the exclusive prefix sum of the cell counts).
__global__ __launch_bounds__(256) void cell_place_kernel(const uint32_t *cell, int64_t n, unsigned long long *cursor,
                                                         uint32_t *order)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const unsigned long long pos = atomicAdd(&cursor[cell[i]], 1ull);
    if (pos < (unsigned long long)n) order[pos] = (uint32_t)i;
}

void launch_multipole_moments(hipStream_t s, const CoilDev *coils, int n_coils, const Tables &tab, double *mp)
{
    if (n_coils > 0)
        hipLaunchKernelGGL(multipole_moments_kernel, dim3(n_coils, kMpChunks), dim3(kMomThreads), 0, s, coils, tab,
                           mp);
}

size_t visit_order_bytes(int64_t n, int bits, size_t scan_bytes)
{
    const size_t cells = (size_t)1 << (3 * bits);
    return sizeof(double) * (6 * kBoxBlocks + 8) + sizeof(uint32_t) * (size_t)n + 2 * sizeof(int64_t) * cells +
           scan_bytes + 6 * 256;
}

hipError_t launch_visit_order(hipStream_t s, const double *pts, int64_t n, int bits, void *temp, size_t scan_bytes,
                              uint32_t *order)
{
    const size_t cells = (size_t)1 << (3 * bits);
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char *t = (char *)temp;
    double *part = (double *)t;
    t += align(sizeof(double) * 6 * kBoxBlocks);
    double *box = (double *)t;
    t += align(sizeof(double) * 8);
    uint32_t *cell = (uint32_t *)t;
    t += align(sizeof(uint32_t) * (size_t)n);
    int64_t *count = (int64_t *)t;
    t += align(sizeof(int64_t) * cells);
    int64_t *cursor = (int64_t *)t;
    t += align(sizeof(int64_t) * cells);
    void *scan_temp = t;
    const int blocks = (int)((n + 255) / 256 < kBoxBlocks ? (n + 255) / 256 : kBoxBlocks);
    hipLaunchKernelGGL(box_partial_kernel, dim3(blocks), dim3(256), 0, s, pts, n, part);
    hipLaunchKernelGGL(box_final_kernel, dim3(1), dim3(256), 0, s, part, blocks, box);
    launch_fill_zero(s, count, (int64_t)cells);
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(cell_count_kernel, dim3(grid), dim3(256), 0, s, pts, n, box, bits, cell,
                       (unsigned long long *)count);
    size_t tb = scan_bytes;
    hipError_t e = near_scan(s, count, cursor, (int64_t)cells, scan_temp, tb);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cell_place_kernel, dim3(grid), dim3(256), 0, s, cell, n, (unsigned long long *)cursor, order);
    return hipGetLastError();
}

}  // namespace ccoil
