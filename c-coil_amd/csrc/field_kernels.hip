// Field kernels for gfx950: all coils of a CoilGroup, all requested quantities, one launch.
//
// Mapping: SPLIT consecutive lanes own one point; the coil loop is wave-uniform, so every coil
// attribute and every Gauss-Legendre table entry is a scalar (SMEM) load broadcast to the wave and
// the method branch (fast/slow) never diverges. Lanes of a point split the thickness rows and are
// reduced with cross-lane shuffles. Results are accumulated in registers over coils (FP64) and each
// output is written once -- the reference instead launches one kernel per coil and read-modify-writes
// the result array each time (FieldBGroup.cu:175-181).
#include "kernels.hpp"
#include "integrands.hpp"
#include "multipole.hpp"

namespace ccoil {

// ---------------------------------------------------------------------------------------------
// Gauss-Legendre tables in FP64 (Newton on P_n), one workgroup per order; the last workgroup writes the far-field
// series' length rule (kSeriesLenNodes nodes) into x and w at kSeriesLenOffset.
// ---------------------------------------------------------------------------------------------
__global__ void gl_table_kernel(double *x, double *w, double *cosv, double *omc, float *xf, float *wf,
                                float *cosf, float *omcf, double *rcos, double *rsin)
{
    static_assert(kSeriesLenNodes <= 128, "one node pair per thread of 64");
    const bool series_len = blockIdx.x == kMaxOrder;
    const int n = series_len ? kSeriesLenNodes : blockIdx.x + 1;
    const int i = threadIdx.x;
    if (i >= (n + 1) / 2) return;
    double z = cos(kPi * (i + 0.75) / (n + 0.5));
    double p0 = 1.0, p1 = z, dp = 1.0;
    for (int it = 0; it < 100; ++it) {
        p0 = 1.0;
        p1 = z;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        if (n == 1) break;
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        double dz = p1 / dp;
        z -= dz;
        if (fabs(dz) < 1e-17) break;
    }
    p0 = 1.0;
    p1 = z;
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    dp = (n == 1) ? 1.0 : n * (z * p1 - p0) / (z * z - 1.0);
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    const int off = series_len ? kSeriesLenOffset : table_offset(n);
    int hi = n - 1 - i, lo = i;
    double xh = z, xl = -z;
    if (hi == lo) { xh = 0.0; xl = 0.0; }
    const int idx[2] = {hi, lo};
    const double xv[2] = {xh, xl};
    for (int k = 0; k < 2; ++k) {
        const int e = off + idx[k];
        const double xx = xv[k];
        x[e] = xx;
        w[e] = wi;
        if (series_len) continue;   // the other tables hold the orders 1..kMaxOrder only
        // cos(pi/2 (1 + x)) = -sin(pi x / 2): exactly antisymmetric in x.
        const double cv = -sinpi(0.5 * xx);
        const double sv = sinpi(0.25 * (1.0 + xx));
        cosv[e] = cv;
        omc[e] = 2.0 * sv * sv;
        xf[e] = (float)xx;
        wf[e] = (float)wi;
        cosf[e] = (float)cv;
        omcf[e] = (float)(2.0 * sv * sv);
        rcos[e] = -cospi(xx);
        rsin[e] = -sinpi(xx);
    }
}

// ---------------------------------------------------------------------------------------------
// Main field kernel.
// ---------------------------------------------------------------------------------------------
template <int SPLIT>
__device__ inline void reduce_local(Local &L)
{
#pragma unroll
    for (int m = 1; m < SPLIT; m <<= 1) {
        L.A += __shfl_xor(L.A, m, SPLIT);
        L.Bh += __shfl_xor(L.Bh, m, SPLIT);
        L.Bz += __shfl_xor(L.Bz, m, SPLIT);
        L.RP += __shfl_xor(L.RP, m, SPLIT);
        L.RR += __shfl_xor(L.RR, m, SPLIT);
        L.RZ += __shfl_xor(L.RZ, m, SPLIT);
        L.ZZ += __shfl_xor(L.ZZ, m, SPLIT);
    }
}

// Minimum waves per SIMD: 5 for the float instances (at most 102 VGPRs; the spills this costs sit at the coil
// level, outside the node loops; the FP32C headline kernel measured 7.6 % faster than at the allocator's 4 waves),
// the allocator's choice for the double instances (forcing more waves there was slower).
template <int PREC> constexpr int kFieldWaves = (PREC == kFP32C || PREC == kFP32) ? 5 : 1;

template <int PREC, bool WA, bool WB, bool WG, int SPLIT>
__global__ __launch_bounds__(kFieldBlock) __attribute__((amdgpu_waves_per_eu(kFieldWaves<PREC>)))
void field_kernel(FieldParams p)
{
    const int64_t t = (int64_t)blockIdx.x * kFieldBlock + threadIdx.x;
    const int64_t vi = t / SPLIT;   // place in the visit order
    const int lane = (int)(t % SPLIT);
    const bool active = vi < p.n_pts;
    const int64_t vc = active ? vi : (p.n_pts - 1);   // inactive lanes shadow a valid point
    const int64_t pi = p.order ? (int64_t)p.order[vc] : vc;
    const double px = p.pts[3 * pi], py = p.pts[3 * pi + 1], pz = p.pts[3 * pi + 2];

    Cart acc;
    for (int k = 0; k < 3; ++k) { acc.A[k] = 0.0; acc.E[k] = 0.0; acc.B[k] = 0.0; }
    for (int k = 0; k < 9; ++k) acc.G[k] = 0.0;

    int n_near = 0;   // FP32C: near-winding coils of this point in this chunk
    const int c0 = blockIdx.y * p.coils_per_chunk, c1 = min(c0 + p.coils_per_chunk, p.n_coils);
    for (int ci = c0; ci < c1; ++ci) {
        const CoilDev &c = p.coils[ci];
        double z, r, cph, sph;
        to_local(c, px, py, pz, z, r, cph, sph);
        Local L = {0, 0, 0, 0, 0, 0, 0};
        if constexpr (PREC == kFP64IEEE) {
            eval_stable<double, WA, WB, WG>(c, p.tab, z, r, lane, SPLIT, L);
        } else if constexpr (PREC == kFP64Reference) {
            eval_plain<double, WA, WB, WG>(c, p.tab, z, r, lane, SPLIT, L);
        } else if constexpr (PREC == kFP32) {
            eval_plain<float, WA, WB, WG>(c, p.tab, z, r, lane, SPLIT, L);
        } else {
            // FP64 and FP32C: beyond the coil's far-field distance (set only when the call uses the series) the FP64
            // series; pairs that become records (FP32C's near shell and axis cone; with rule records every other
            // pair) nothing here: the record pass adds them; elsewhere the rule
            const bool far = is_far(c, z, r);
            const int kind = record_kind(p, c, z, r, WG);
            if (!far && kind < 0) {
                if constexpr (PREC == kFP64) eval_stable<f64s, WA, WB, WG>(c, p.tab, z, r, lane, SPLIT, L);
                else eval_stable<float, WA, WB, WG>(c, p.tab, z, r, lane, SPLIT, L);
            }
            if (kind >= 0) ++n_near;
            if (__ballot(far)) {
                // the loop runs to the wave's largest term count (uniform bounds and coefficient loads); each lane
                // stops adding at its own count, so its value does not depend on the other points of its wave
                const int own = far ? multipole_terms(r * r + z * z, c.inv_rmax2,
                                                      PREC == kFP64 ? kMpLogTolFP64 : kMpLogTolFP32C) : 0;
                const int terms = wave_max_small(own), lo = wave_min_small(own, far);
                if (far && lane == 0)
                    eval_multipole<WA, WB, WG>(p.mp + kMpHeader + (size_t)ci * kMpCoil, p.mp, c.rmax, z, r, terms,
                                               lo, own, L);
            }
        }
        if (SPLIT > 1) reduce_local<SPLIT>(L);
        accumulate_global(c, L, r, cph, sph, p.q, acc);
    }
    if (!active || lane != 0) return;
    if ((PREC == kFP32C || PREC == kFP64) && n_near > 0)   // integer sums: exact in any order
        atomicAdd((unsigned long long *)&p.near_counts[pi], (unsigned long long)n_near);
    if (p.n_chunks > 1) {
        double *o = p.partial + ((int64_t)blockIdx.y * p.n_pts + pi) * p.partial_width;
        if (p.A) { for (int k = 0; k < 3; ++k) o[k] = acc.A[k]; o += 3; }
        if (p.E) { for (int k = 0; k < 3; ++k) o[k] = acc.E[k]; o += 3; }
        if (p.B) { for (int k = 0; k < 3; ++k) o[k] = acc.B[k]; o += 3; }
        if (p.G) for (int k = 0; k < 9; ++k) o[k] = acc.G[k];
        return;
    }
    if (p.A) for (int k = 0; k < 3; ++k) p.A[3 * pi + k] = acc.A[k];
    if (p.E) for (int k = 0; k < 3; ++k) p.E[3 * pi + k] = acc.E[k];
    if (p.B) for (int k = 0; k < 3; ++k) p.B[3 * pi + k] = acc.B[k];
    if (p.G) for (int k = 0; k < 9; ++k) p.G[9 * pi + k] = acc.G[k];
}

// The series alone, for calls whose record pass takes every pair inside the coils' far-field distances (FP64 or FP32C
// with rule records): one lane per point, the coils looped inside; a pair beyond the distance takes the series, a
// pair inside it is counted for the record pass. Without the rule's node loops the kernel holds far fewer registers
// than field_kernel, so more waves hide the latency of the coefficient loads. A point's value is its far coils'
// series in coil order, the same sums as field_kernel's. It also counts the records of each coil key, which
// near_count_kernel does for the other calls with a near pass.
template <int PREC, bool WA, bool WB, bool WG>
__global__ __launch_bounds__(kFieldBlock) __attribute__((amdgpu_waves_per_eu(4)))
void field_series_kernel(FieldParams p)
{
    const int64_t vi = (int64_t)blockIdx.x * kFieldBlock + threadIdx.x;
    const bool active = vi < p.n_pts;
    const int64_t vc = active ? vi : (p.n_pts - 1);   // inactive lanes shadow a valid point
    const int64_t pi = p.order ? (int64_t)p.order[vc] : vc;
    const double px = p.pts[3 * pi], py = p.pts[3 * pi + 1], pz = p.pts[3 * pi + 2];

    Cart acc;
    for (int k = 0; k < 3; ++k) { acc.A[k] = 0.0; acc.E[k] = 0.0; acc.B[k] = 0.0; }
    for (int k = 0; k < 9; ++k) acc.G[k] = 0.0;

    int n_near = 0;
    const int lane = threadIdx.x & 63;
    const int c0 = blockIdx.y * p.coils_per_chunk, c1 = min(c0 + p.coils_per_chunk, p.n_coils);
    // the record pass's per-key counts (its cursors come from their scan), per block in LDS: one ballot and one LDS
    // atomic per wave, coil and kind in the loop, one global atomic per nonzero key after it. A global write inside
    // the loop would keep the compiler from loading the coil records through the scalar cache.
    extern __shared__ uint32_t key_counts[];   // [kind][coil of the chunk]
    const bool count = series_counts_keys(p);
    const int n_kinds = p.float_records ? 2 : 1, cpc = p.coils_per_chunk;
    if (count)
        for (int i = threadIdx.x; i < n_kinds * cpc; i += kFieldBlock) key_counts[i] = 0u;
    __syncthreads();
    for (int ci = c0; ci < c1; ++ci) {
        const CoilDev &c = p.coils[ci];
        double z, r, cph, sph;
        to_local(c, px, py, pz, z, r, cph, sph);
        const bool far = is_far(c, z, r);
        const int kind = active ? record_kind(p, c, z, r, WG) : -1;
        if (kind >= 0) ++n_near;
        if (count)
            for (int k = 0; k < n_kinds; ++k) {
                const uint64_t m = __ballot(kind == k);
                if (m && lane == __builtin_ctzll(m)) atomicAdd(&key_counts[k * cpc + (ci - c0)], (unsigned)__popcll(m));
            }
        if (__ballot(far)) {
            const int own = far ? multipole_terms(r * r + z * z, c.inv_rmax2,
                                                  PREC == kFP64 ? kMpLogTolFP64 : kMpLogTolFP32C) : 0;
            const int terms = wave_max_small(own), lo = wave_min_small(own, far);
            if (far) {
                Local L;
                eval_multipole<WA, WB, WG>(p.mp + kMpHeader + (size_t)ci * kMpCoil, p.mp, c.rmax, z, r, terms, lo,
                                           own, L);
                accumulate_global(c, L, r, cph, sph, p.q, acc);
            }
        }
    }
    __syncthreads();
    if (count)
        for (int i = threadIdx.x; i < n_kinds * cpc; i += kFieldBlock) {
            const uint32_t v = key_counts[i];
            const int k = i / cpc, ci = c0 + i % cpc;
            if (v && ci < c1) atomicAdd(&p.near_hist[k * p.n_coils + p.coils[ci].near_key], v);
        }
    if (!active) return;
    if (n_near > 0)   // integer sums: exact in any order
        atomicAdd((unsigned long long *)&p.near_counts[pi], (unsigned long long)n_near);
    if (p.n_chunks > 1) {
        double *o = p.partial + ((int64_t)blockIdx.y * p.n_pts + pi) * p.partial_width;
        if (p.A) { for (int k = 0; k < 3; ++k) o[k] = acc.A[k]; o += 3; }
        if (p.E) { for (int k = 0; k < 3; ++k) o[k] = acc.E[k]; o += 3; }
        if (p.B) { for (int k = 0; k < 3; ++k) o[k] = acc.B[k]; o += 3; }
        if (p.G) for (int k = 0; k < 9; ++k) o[k] = acc.G[k];
        return;
    }
    if (p.A) for (int k = 0; k < 3; ++k) p.A[3 * pi + k] = acc.A[k];
    if (p.E) for (int k = 0; k < 3; ++k) p.E[3 * pi + k] = acc.E[k];
    if (p.B) for (int k = 0; k < 3; ++k) p.B[3 * pi + k] = acc.B[k];
    if (p.G) for (int k = 0; k < 9; ++k) p.G[9 * pi + k] = acc.G[k];
}

// Sum of the coil chunks, in chunk order (deterministic), into the requested outputs. One thread per
// (point, partial component), so consecutive lanes read consecutive partial words and write consecutive output
// words.
__global__ __launch_bounds__(256) void field_reduce_kernel(FieldParams p)
{
    const int w = p.partial_width;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= p.n_pts * w) return;
    double s = 0.0;
    for (int ch = 0; ch < p.n_chunks; ++ch) s += p.partial[(int64_t)ch * p.n_pts * w + t];
    const int64_t pi = t / w;
    int k = (int)(t - pi * w);
    if (p.A) { if (k < 3) { p.A[3 * pi + k] = s; return; } k -= 3; }
    if (p.E) { if (k < 3) { p.E[3 * pi + k] = s; return; } k -= 3; }
    if (p.B) { if (k < 3) { p.B[3 * pi + k] = s; return; } k -= 3; }
    if (p.G) p.G[9 * pi + k] = s;
}

// ---------------------------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------------------------
void launch_gl_tables(hipStream_t s, double *x, double *w, double *cosv, double *omc, float *xf, float *wf,
                      float *cosf, float *omcf, double *rcos, double *rsin)
{
    hipLaunchKernelGGL(gl_table_kernel, dim3(kMaxOrder + 1), dim3(64), 0, s, x, w, cosv, omc, xf, wf, cosf, omcf,
                       rcos, rsin);
}

// Derived angular tables: w cos(pi/2 (1 + x)) and 1 + cos(pi/2 (1 + x)) = 2 cos^2(pi/4 (1 + x)) (no cancellation
// near phi = pi).
__global__ void angular_table_kernel(const double *x, const double *w, const double *cosv, const double *omc,
                                     double *wc, float *wcf, double *opc, float *opcf, double *ws, double *ws2,
                                     double *wcs, float *wcsf)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= kTableSize) return;
    wc[e] = w[e] * cosv[e];
    wcf[e] = (float)wc[e];
    const double c = cospi(0.25 * (1.0 + x[e]));
    opc[e] = 2.0 * c * c;
    opcf[e] = (float)opc[e];
    ws[e] = w[e] * omc[e];
    ws2[e] = ws[e] * omc[e];
    wcs[e] = wc[e] * omc[e];
    wcsf[e] = (float)wcs[e];
}

void launch_angular_tables(hipStream_t s, const double *x, const double *w, const double *cosv, const double *omc,
                           double *wc, float *wcf, double *opc, float *opcf, double *ws, double *ws2, double *wcs,
                           float *wcsf)
{
    hipLaunchKernelGGL(angular_table_kernel, dim3((kTableSize + 255) / 256), dim3(256), 0, s, x, w, cosv, omc, wc,
                       wcf, opc, opcf, ws, ws2, wcs, wcsf);
}

__global__ void fill_zero_kernel(unsigned long long *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

void launch_fill_zero(hipStream_t s, void *p, int64_t n_words)
{
    if (n_words <= 0) return;
    const unsigned blocks = (unsigned)((n_words + 255) / 256);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, s, (unsigned long long *)p, n_words);
}

template <int PREC, bool WA, bool WB, bool WG>
static void launch_field_q(hipStream_t s, const FieldParams &p, int split)
{
    const int64_t threads = p.n_pts * split;
    const dim3 grid((unsigned)((threads + kFieldBlock - 1) / kFieldBlock), (unsigned)p.n_chunks);
    if (split == 1) hipLaunchKernelGGL((field_kernel<PREC, WA, WB, WG, 1>), grid, dim3(kFieldBlock), 0, s, p);
    else if (split == 4) hipLaunchKernelGGL((field_kernel<PREC, WA, WB, WG, 4>), grid, dim3(kFieldBlock), 0, s, p);
    else hipLaunchKernelGGL((field_kernel<PREC, WA, WB, WG, 16>), grid, dim3(kFieldBlock), 0, s, p);
}

template <int PREC, bool WA, bool WB, bool WG>
static void launch_series_q(hipStream_t s, const FieldParams &p)
{
    const dim3 grid((unsigned)((p.n_pts + kFieldBlock - 1) / kFieldBlock), (unsigned)p.n_chunks);
    const size_t lds = series_counts_keys(p) ? sizeof(uint32_t) * (p.float_records ? 2 : 1) * p.coils_per_chunk : 0;
    hipLaunchKernelGGL((field_series_kernel<PREC, WA, WB, WG>), grid, dim3(kFieldBlock), lds, s, p);
}

template <int PREC>
static void launch_series_p(hipStream_t s, const FieldParams &p)
{
    const bool wa = p.q & (Q_A | Q_E), wb = p.q & Q_B, wg = p.q & (Q_G | Q_GJ);
    switch ((wa ? 4 : 0) | (wb ? 2 : 0) | (wg ? 1 : 0)) {
        case 1: launch_series_q<PREC, false, false, true>(s, p); break;
        case 2: launch_series_q<PREC, false, true, false>(s, p); break;
        case 3: launch_series_q<PREC, false, true, true>(s, p); break;
        case 4: launch_series_q<PREC, true, false, false>(s, p); break;
        case 5: launch_series_q<PREC, true, false, true>(s, p); break;
        case 6: launch_series_q<PREC, true, true, false>(s, p); break;
        default: launch_series_q<PREC, true, true, true>(s, p); break;
    }
}

template <int PREC>
static void launch_field_p(hipStream_t s, const FieldParams &p, int split)
{
    if constexpr (PREC == kFP64 || PREC == kFP32C) {
        if (p.rule_records) {   // the record pass takes every pair inside the far-field distances
            launch_series_p<PREC>(s, p);
            return;
        }
    }
    const bool wa = p.q & (Q_A | Q_E), wb = p.q & Q_B, wg = p.q & (Q_G | Q_GJ);
    const int code = (wa ? 4 : 0) | (wb ? 2 : 0) | (wg ? 1 : 0);
    switch (code) {
        case 1: launch_field_q<PREC, false, false, true>(s, p, split); break;
        case 2: launch_field_q<PREC, false, true, false>(s, p, split); break;
        case 3: launch_field_q<PREC, false, true, true>(s, p, split); break;
        case 4: launch_field_q<PREC, true, false, false>(s, p, split); break;
        case 5: launch_field_q<PREC, true, false, true>(s, p, split); break;
        case 6: launch_field_q<PREC, true, true, false>(s, p, split); break;
        default: launch_field_q<PREC, true, true, true>(s, p, split); break;
    }
}

template <int PREC, bool WA, bool WB, bool WG>
static int blocks_per_cu_q(int split)
{
    int b = 0;
    hipError_t e =
        split == 1   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, field_kernel<PREC, WA, WB, WG, 1>, kFieldBlock, 0)
        : split == 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, field_kernel<PREC, WA, WB, WG, 4>, kFieldBlock, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, field_kernel<PREC, WA, WB, WG, 16>, kFieldBlock, 0);
    return e == hipSuccess ? b : 0;
}

template <int PREC>
static int blocks_per_cu_p(uint32_t q, int split)
{
    const bool wa = q & (Q_A | Q_E), wb = q & Q_B, wg = q & (Q_G | Q_GJ);
    switch ((wa ? 4 : 0) | (wb ? 2 : 0) | (wg ? 1 : 0)) {
        case 1: return blocks_per_cu_q<PREC, false, false, true>(split);
        case 2: return blocks_per_cu_q<PREC, false, true, false>(split);
        case 3: return blocks_per_cu_q<PREC, false, true, true>(split);
        case 4: return blocks_per_cu_q<PREC, true, false, false>(split);
        case 5: return blocks_per_cu_q<PREC, true, false, true>(split);
        case 6: return blocks_per_cu_q<PREC, true, true, false>(split);
        default: return blocks_per_cu_q<PREC, true, true, true>(split);
    }
}

template <int PREC, bool WA, bool WB, bool WG>
static int series_blocks_q()
{
    int b = 0;
    const hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, field_series_kernel<PREC, WA, WB, WG>, kFieldBlock, 0);
    return e == hipSuccess ? b : 0;
}

template <int PREC>
static int series_blocks_per_cu_p(uint32_t q)
{
    const bool wa = q & (Q_A | Q_E), wb = q & Q_B, wg = q & (Q_G | Q_GJ);
    switch ((wa ? 4 : 0) | (wb ? 2 : 0) | (wg ? 1 : 0)) {
        case 1: return series_blocks_q<PREC, false, false, true>();
        case 2: return series_blocks_q<PREC, false, true, false>();
        case 3: return series_blocks_q<PREC, false, true, true>();
        case 4: return series_blocks_q<PREC, true, false, false>();
        case 5: return series_blocks_q<PREC, true, false, true>();
        case 6: return series_blocks_q<PREC, true, true, false>();
        default: return series_blocks_q<PREC, true, true, true>();
    }
}

int field_series_blocks_per_cu(uint32_t q, int precision)
{
    return precision == kFP64 ? series_blocks_per_cu_p<kFP64>(q) : series_blocks_per_cu_p<kFP32C>(q);
}

int field_blocks_per_cu(uint32_t q, int precision, int split)
{
    if (precision == kFP64) return blocks_per_cu_p<kFP64>(q, split);
    if (precision == kFP64IEEE) return blocks_per_cu_p<kFP64IEEE>(q, split);
    if (precision == kFP64Reference) return blocks_per_cu_p<kFP64Reference>(q, split);
    if (precision == kFP32) return blocks_per_cu_p<kFP32>(q, split);
    return blocks_per_cu_p<kFP32C>(q, split);
}

void launch_field(hipStream_t s, const FieldParams &p, int precision, int split)
{
    if (p.n_pts <= 0) return;
    if (precision == kFP64) launch_field_p<kFP64>(s, p, split);
    else if (precision == kFP64IEEE) launch_field_p<kFP64IEEE>(s, p, split);
    else if (precision == kFP64Reference) launch_field_p<kFP64Reference>(s, p, split);
    else if (precision == kFP32) launch_field_p<kFP32>(s, p, split);
    else launch_field_p<kFP32C>(s, p, split);
}

void launch_field_reduce(hipStream_t s, const FieldParams &p)
{
    if (p.n_pts <= 0 || p.n_chunks <= 1) return;
    const int64_t n = p.n_pts * p.partial_width;
    hipLaunchKernelGGL(field_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p);
}

}  // namespace ccoil
