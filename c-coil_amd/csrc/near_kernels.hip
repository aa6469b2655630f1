// Record pass (gfx950): the (coil, point) pairs field_kernel defers. In FP32C those inside a coil's near shell (and,
// for outputs that divide by r, its axis cone) are re-evaluated in FP64; with rule records (the series modes, large
// calls) every pair inside the far-field distance is, by the FP64 or the float rule (its record's kind), on waves
// that share one coil instead of beside the series.
//
// field_kernel<kFP32C> skips those pairs and counts them per point; an exclusive scan of the point counts gives each
// point's slots; near_count_kernel counts each coil's near points and near_key_scan_kernel turns those counts into
// cursors (a counting sort by coil) and checks the total against the buffers' capacity; near_fill_kernel places one
// (coil key, point | slot) record per pair at its coil's cursor, so that the lanes of a wave evaluate the same coil
// (uniform loop bounds, one branch of the fast / slow rule); near_eval_kernel evaluates one record per thread into
// its slot; near_add_kernel adds a point's slots to the outputs in coil order. A slot's value depends only on its
// (coil, point), and every sum has a fixed order, so FP32C results are bitwise reproducible although the order of the
// records within a coil is not. Nothing here is read back by the host before the work is done.
#include <cstddef>

#include "kernels.hpp"
#include "integrands.hpp"

namespace ccoil {

// Whether the points of a wave may be records of the coils c0 .. c0 + 7 (bit j for coil c0 + j < n_coils): false
// only where every point lies beyond the coil's far-field sphere by more than a margin, tested in the global frame
// (the rotation to the coil's frame keeps lengths up to a few ulps, far below the margin). The exact class is
// record_kind's; this lets a wave skip the frame transform of the coils none of its points can be a record of. The
// coils' centres and far_d2 are read through the constant address space: the records do not change during a
// launch, so these are scalar loads at wave-uniform indices although the loops that call this also store to global
// memory (through a generic pointer their loads would be vector loads, each a full memory round trip). The eight
// coils' loads come first, so a wave waits for them once.
__device__ inline uint32_t record_candidates(const CoilDev *coils, int n_coils, int c0, double px, double py,
                                             double pz)
{
    typedef const __attribute__((address_space(4))) double cdouble;
    constexpr int kPos = offsetof(CoilDev, pos) / sizeof(double), kFar = offsetof(CoilDev, far_d2) / sizeof(double);
    double fd[8], cx[8], cy[8], cz[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {   // past the last coil, the last coil again (its bit is cleared below)
        cdouble *f = (cdouble *)(const void *)(coils + min(c0 + j, n_coils - 1));
        fd[j] = f[kFar];
        cx[j] = f[kPos];
        cy[j] = f[kPos + 1];
        cz[j] = f[kPos + 2];
    }
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const double dx = px - cx[j], dy = py - cy[j], dz = pz - cz[j];
        const bool may = !(fd[j] > 0.0) | (dx * dx + dy * dy + dz * dz <= fd[j] * (1.0 + 1e-9));
        m |= (__ballot(may) != 0 && c0 + j < n_coils) ? 1u << j : 0u;
    }
    return m;
}

// Calls with fewer points than kNearSplitPoints split each near record's thickness rows over kNearSplit lanes.
constexpr int64_t kNearSplitPoints = (int64_t)1 << 17;
constexpr int kNearSplit = 8;

// Per coil key, the near points (points with near pairs only; the coil loop is wave-uniform, so each coil's count
// in a wave is one ballot and one atomic).
__global__ __launch_bounds__(256) void near_count_kernel(FieldParams p)
{
    const int64_t vi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (vi >= p.n_pts) return;
    const int64_t pi = p.order ? (int64_t)p.order[vi] : vi;   // the field kernel's visit order
    if (p.near_counts[pi] == 0) return;
    const double px = p.pts[3 * pi], py = p.pts[3 * pi + 1], pz = p.pts[3 * pi + 2];
    const int lane = threadIdx.x & 63;
    // coils in ascending order, skipping those no point of the wave can be a record of
    for (int c0 = 0; c0 < p.n_coils; c0 += 8)
        for (uint32_t cand = record_candidates(p.coils, p.n_coils, c0, px, py, pz); cand; cand &= cand - 1) {
            const CoilDev &c = p.coils[c0 + __builtin_ctz(cand)];
            double z, r, cph, sph;
            to_local(c, px, py, pz, z, r, cph, sph);
            const int kind = record_kind(p, c, z, r, p.q & (Q_G | Q_GJ));   // the field kernel's test, bitwise
            for (int k = 0; k <= p.float_records; ++k) {
                const uint64_t m = __ballot(kind == k);
                if (m && lane == __builtin_ctzll(m))
                    atomicAdd(&p.near_hist[k * p.n_coils + c.near_key], (unsigned)__popcll(m));
            }
        }
}

// One workgroup: cursor = exclusive prefix sum of near_hist over the n_coils + 1 keys, and the call's total.
__global__ __launch_bounds__(256) void near_key_scan_kernel(FieldParams p)
{
    __shared__ uint32_t part[256];
    const int t = threadIdx.x, nk = p.n_keys + 1;
    uint32_t carry = 0;
    for (int base = 0; base < nk; base += 256) {
        const uint32_t v = base + t < nk ? p.near_hist[base + t] : 0u;
        part[t] = v;
        __syncthreads();
        for (int d = 1; d < 256; d <<= 1) {   // inclusive Hillis-Steele scan
            const uint32_t add = t >= d ? part[t - d] : 0u;
            __syncthreads();
            part[t] += add;
            __syncthreads();
        }
        if (base + t < nk) p.near_cursor[base + t] = carry + part[t] - v;
        carry += part[255];
        __syncthreads();
    }
    if (t == 0) {
        const unsigned long long total =
            (unsigned long long)(p.near_offsets[p.n_pts - 1] + p.near_counts[p.n_pts - 1]);
        p.near_state->total = total;
        p.near_state->overflow = total > (unsigned long long)p.near_cap ? 1u : 0u;
        p.near_state->fp64 = p.near_cursor[p.n_coils];   // the kind-0 keys come first
    }
}

__global__ __launch_bounds__(256) void near_fill_kernel(FieldParams p)
{
    if (p.near_state->overflow) return;
    const int64_t vi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (vi >= p.n_pts) return;
    const int64_t pi = p.order ? (int64_t)p.order[vi] : vi;
    const int64_t count = p.near_counts[pi];
    if (count == 0) return;
    const uint64_t total = p.near_state->total;
    const double px = p.pts[3 * pi], py = p.pts[3 * pi + 1], pz = p.pts[3 * pi + 2];
    int64_t slot = p.near_offsets[pi];
    const int64_t end = slot + count;
    const uint64_t hi = (uint64_t)pi << 32;
    const int lane = threadIdx.x & 63;
    bool bad = false;
    // coils in ascending order (a point's slots follow its coils), skipping those no point of the wave can be a
    // record of
    for (int c0 = 0; c0 < p.n_coils; c0 += 8) {
        for (uint32_t cand = record_candidates(p.coils, p.n_coils, c0, px, py, pz); cand; cand &= cand - 1) {
            const CoilDev &c = p.coils[c0 + __builtin_ctz(cand)];
            double z, r, cph, sph;
            to_local(c, px, py, pz, z, r, cph, sph);
            const int kind = record_kind(p, c, z, r, p.q & (Q_G | Q_GJ));   // the field kernel's test, bitwise
            // the wave's records of this coil and kind take consecutive places from one atomic on the key's cursor
            for (int k = 0; k <= p.float_records; ++k) {
                const bool mine = kind == k;
                const uint64_t m = __ballot(mine);
                if (!m) continue;
                const int leader = __builtin_ctzll(m);
                const uint32_t key = (uint32_t)(k * p.n_coils + c.near_key);
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&p.near_cursor[key], (unsigned)__popcll(m));
                base = __shfl(base, leader);
                if (!mine) continue;
                const uint64_t pos = base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
                // a pair classified differently from the series or field kernel: report it, write nothing
                if (slot >= end || pos >= total) {
                    bad = true;
                    continue;
                }
                p.near_keys[pos] = key;
                p.near_recs[pos] = hi | (uint64_t)slot;
                ++slot;
            }
        }
    }
    if (bad || slot != end) atomicOr(&p.near_state->mismatch, 1u);
}

// Records sorted by coil: S consecutive lanes take one record (thickness rows sub, sub + S, ..., combined by a
// fixed xor butterfly, so every lane of the group holds the same sum); its value goes to the record's slot. S > 1 on
// calls with few points, where the records are too few to fill the chip one lane each and one lane's serial chain
// of nodes sets the kernel's duration; the lanes then stride over the records (a group's value does not depend on
// the grid). FLOAT: the kind-1 records [fp64, total) by the float rule, at 5 waves per SIMD as field_kernel's float
// instances; else the kind-0 records [0, fp64) by the FP64 rule.
template <bool WA, bool WB, bool WG, bool FLOAT, int S, bool LOOP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FLOAT ? 5 : 1)))
void near_eval_kernel(FieldParams p)
{
    if (p.near_state->overflow) return;
    const int64_t total = (int64_t)p.near_state->total, fp64 = (int64_t)p.near_state->fp64;
    const int64_t lo = FLOAT ? fp64 : 0, hi = FLOAT ? total : fp64;
    const int sub = (int)(threadIdx.x % S);
    auto record = [&](int64_t e) {
        const uint32_t key = p.near_keys[e];
        const uint64_t rec = p.near_recs[e];
        const int64_t pi = (int64_t)(rec >> 32), slot = (int64_t)(rec & 0xffffffffu);
        // only a near_fill that reported a mismatch leaves records unwritten; never follow one out of bounds
        if (pi >= p.n_pts || slot >= total) return;
        double *o = p.near_vals + slot * p.near_width;
        // A wave holds one coil's records, or two or more at the boundaries of the sorted keys: it takes its keys
        // one at a time (the key of the lowest lane still to do, read into a scalar register), and every lane
        // evaluates that coil at its point, keeping the value only if the key is its own. The control flow stays
        // uniform and no store precedes the loads, so the coil's fields and the quadrature tables are scalar loads,
        // as in field_kernel; a divergent branch would take the same issue time. todo is wave-uniform and loses at
        // least one lane on every pass; a group's lanes share their key, so a group is always active as a whole.
        const double px = p.pts[3 * pi], py = p.pts[3 * pi + 1], pz = p.pts[3 * pi + 2];
        Cart acc;
        for (int j = 0; j < 3; ++j) { acc.A[j] = 0.0; acc.E[j] = 0.0; acc.B[j] = 0.0; }
        for (int j = 0; j < 9; ++j) acc.G[j] = 0.0;
        uint64_t todo = __ballot(1);
        while (todo) {
            const uint32_t k = __builtin_amdgcn_readfirstlane(__shfl(key, __builtin_ctzll(todo)));
            const bool mine = key == k;
            todo &= ~(uint64_t)__ballot(mine);
            // only a reported mismatch can leave a key outside the range's kind
            const uint32_t kind = k / (uint32_t)p.n_coils;
            if (k >= (uint32_t)p.n_keys || kind != (FLOAT ? 1u : 0u)) continue;
            const CoilDev &c = p.coils[p.near_coil[k - kind * (uint32_t)p.n_coils]];
            double z, r, cph, sph;
            to_local(c, px, py, pz, z, r, cph, sph);
            Local L;
            if (FLOAT) eval_stable<float, WA, WB, WG>(c, p.tab, z, r, sub, S, L);
            else eval_stable<f64s, WA, WB, WG>(c, p.tab, z, r, sub, S, L);
            if constexpr (S > 1) {
#pragma unroll
                for (int m = 1; m < S; m <<= 1) {
                    L.A += __shfl_xor(L.A, m, 64);
                    L.Bh += __shfl_xor(L.Bh, m, 64);
                    L.Bz += __shfl_xor(L.Bz, m, 64);
                    L.RP += __shfl_xor(L.RP, m, 64);
                    L.RR += __shfl_xor(L.RR, m, 64);
                    L.RZ += __shfl_xor(L.RZ, m, 64);
                    L.ZZ += __shfl_xor(L.ZZ, m, 64);
                }
            }
            if (mine) accumulate_global(c, L, r, cph, sph, p.q, acc);
        }
        if (sub == 0) {
            if (p.A) { for (int j = 0; j < 3; ++j) o[j] = acc.A[j]; o += 3; }
            if (p.E) { for (int j = 0; j < 3; ++j) o[j] = acc.E[j]; o += 3; }
            if (p.B) { for (int j = 0; j < 3; ++j) o[j] = acc.B[j]; o += 3; }
            if (p.G) for (int j = 0; j < 9; ++j) o[j] = acc.G[j];
        }
    };
    const int64_t e0 = lo + ((int64_t)blockIdx.x * 256 + threadIdx.x) / S;
    if constexpr (!LOOP) {
        // the grid covers the capacity, one record per group; with no store before the loads, the coil records
        // and the tables are scalar loads
        if (e0 < hi) record(e0);
    } else {
        // a bounded grid strides over a larger capacity (the stores in the loop make the loads vector loads). A
        // wave's groups hold consecutive records, so the loop's exit is wave-uniform once no lane has a record
        // left, and a group is active or idle as a whole
        const int64_t stride = (int64_t)gridDim.x * (256 / S);
        for (int64_t e = e0; __ballot(e < hi); e += stride)
            if (e < hi) record(e);
    }
}

__global__ __launch_bounds__(256) void near_add_kernel(FieldParams p)
{
    if (p.near_state->overflow) return;
    const int64_t vi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (vi >= p.n_pts) return;
    const int64_t pi = p.order ? (int64_t)p.order[vi] : vi;
    const int64_t count = p.near_counts[pi];
    if (count == 0) return;
    const int64_t e0 = p.near_offsets[pi];
    // the point's outputs are held in registers and each slot added in turn: the same order of additions as a
    // read-modify-write per slot, with one load and one store per output
    double va[3] = {}, ve[3] = {}, vb[3] = {}, vg[9] = {};
    if (p.A) for (int k = 0; k < 3; ++k) va[k] = p.A[3 * pi + k];
    if (p.E) for (int k = 0; k < 3; ++k) ve[k] = p.E[3 * pi + k];
    if (p.B) for (int k = 0; k < 3; ++k) vb[k] = p.B[3 * pi + k];
    if (p.G) for (int k = 0; k < 9; ++k) vg[k] = p.G[9 * pi + k];
    for (int64_t e = e0; e < e0 + count; ++e) {
        const double *o = p.near_vals + e * p.near_width;
        if (p.A) { for (int k = 0; k < 3; ++k) va[k] += o[k]; o += 3; }
        if (p.E) { for (int k = 0; k < 3; ++k) ve[k] += o[k]; o += 3; }
        if (p.B) { for (int k = 0; k < 3; ++k) vb[k] += o[k]; o += 3; }
        if (p.G) for (int k = 0; k < 9; ++k) vg[k] += o[k];
    }
    if (p.A) for (int k = 0; k < 3; ++k) p.A[3 * pi + k] = va[k];
    if (p.E) for (int k = 0; k < 3; ++k) p.E[3 * pi + k] = ve[k];
    if (p.B) for (int k = 0; k < 3; ++k) p.B[3 * pi + k] = vb[k];
    if (p.G) for (int k = 0; k < 9; ++k) p.G[9 * pi + k] = vg[k];
}

// Exclusive prefix sum of the per-point counts in three kernels (tile sums, a one-workgroup scan of the tile sums,
// the tiles' own scans): only kernel launches, so a caller may capture it in a graph, and deterministic.
constexpr int kScanPerThread = 16;
constexpr int kScanTile = 256 * kScanPerThread;

__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t *lds, int64_t &total)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int64_t add = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    const int64_t incl = lds[t];
    total = lds[255];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(256) void scan_tile_sums_kernel(const int64_t *counts, int64_t n, int64_t *sums)
{
    __shared__ int64_t lds[256];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    int64_t v = 0;
    for (int k = 0; k < kScanPerThread; ++k)
        if (base + k < n) v += counts[base + k];
    int64_t total;
    (void)block_exclusive_scan(v, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_sums_kernel(int64_t *sums, int64_t nb)
{
    __shared__ int64_t lds[256];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += 256) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < nb ? sums[i] : 0;
        int64_t total;
        const int64_t ex = block_exclusive_scan(v, lds, total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(256) void scan_tiles_kernel(const int64_t *counts, int64_t n, const int64_t *sums,
                                                         int64_t *offsets)
{
    __shared__ int64_t lds[256];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    int64_t v[kScanPerThread], t = 0;
    for (int k = 0; k < kScanPerThread; ++k) {
        v[k] = base + k < n ? counts[base + k] : 0;
        t += v[k];
    }
    int64_t total;
    int64_t run = sums[blockIdx.x] + block_exclusive_scan(t, lds, total);
    for (int k = 0; k < kScanPerThread; ++k) {
        if (base + k < n) offsets[base + k] = run;
        run += v[k];
    }
}

hipError_t near_scan(hipStream_t s, const int64_t *counts, int64_t *offsets, int64_t n, void *temp,
                     size_t &temp_bytes)
{
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (!temp) {
        temp_bytes = sizeof(int64_t) * (size_t)(nb > 0 ? nb : 1);
        return hipSuccess;
    }
    if (n <= 0) return hipSuccess;
    int64_t *sums = (int64_t *)temp;
    hipLaunchKernelGGL(scan_tile_sums_kernel, dim3((unsigned)nb), dim3(256), 0, s, counts, n, sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, s, sums, nb);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)nb), dim3(256), 0, s, counts, n, (const int64_t *)sums,
                       offsets);
    return hipGetLastError();
}

void launch_near_key_scan(hipStream_t s, const FieldParams &p)
{
    if (p.n_pts <= 0) return;
    hipLaunchKernelGGL(near_key_scan_kernel, dim3(1), dim3(256), 0, s, p);
}

void launch_near_count(hipStream_t s, const FieldParams &p)
{
    if (p.n_pts <= 0) return;
    hipLaunchKernelGGL(near_count_kernel, dim3((unsigned)((p.n_pts + 255) / 256)), dim3(256), 0, s, p);
}

void launch_near_fill(hipStream_t s, const FieldParams &p)
{
    if (p.n_pts <= 0) return;
    hipLaunchKernelGGL(near_fill_kernel, dim3((unsigned)((p.n_pts + 255) / 256)), dim3(256), 0, s, p);
}

template <bool FLOAT, int S, bool LOOP>
static void launch_near_eval_g(hipStream_t s, const FieldParams &p, unsigned blocks)
{
    const bool wa = p.q & (Q_A | Q_E), wb = p.q & Q_B, wg = p.q & (Q_G | Q_GJ);
    const dim3 g(blocks), b(256);
    switch ((wa ? 4 : 0) | (wb ? 2 : 0) | (wg ? 1 : 0)) {
        case 1: hipLaunchKernelGGL((near_eval_kernel<false, false, true, FLOAT, S, LOOP>), g, b, 0, s, p); break;
        case 2: hipLaunchKernelGGL((near_eval_kernel<false, true, false, FLOAT, S, LOOP>), g, b, 0, s, p); break;
        case 3: hipLaunchKernelGGL((near_eval_kernel<false, true, true, FLOAT, S, LOOP>), g, b, 0, s, p); break;
        case 4: hipLaunchKernelGGL((near_eval_kernel<true, false, false, FLOAT, S, LOOP>), g, b, 0, s, p); break;
        case 5: hipLaunchKernelGGL((near_eval_kernel<true, false, true, FLOAT, S, LOOP>), g, b, 0, s, p); break;
        case 6: hipLaunchKernelGGL((near_eval_kernel<true, true, false, FLOAT, S, LOOP>), g, b, 0, s, p); break;
        default: hipLaunchKernelGGL((near_eval_kernel<true, true, true, FLOAT, S, LOOP>), g, b, 0, s, p); break;
    }
}

// S = 1: one lane per record of the capacity. S > 1: S lanes per record of the capacity while that is at most 2^21
// lanes, else a grid of 2^21 lanes striding over the records.
template <bool FLOAT, int S>
static void launch_near_eval_k(hipStream_t s, const FieldParams &p)
{
    const int64_t lanes = p.near_cap * S, max_lanes = (int64_t)1 << 21;
    if constexpr (S == 1) {
        launch_near_eval_g<FLOAT, 1, false>(s, p, (unsigned)((lanes + 255) / 256));
    } else {
        if (lanes <= max_lanes) launch_near_eval_g<FLOAT, S, false>(s, p, (unsigned)((lanes + 255) / 256));
        else launch_near_eval_g<FLOAT, S, true>(s, p, (unsigned)(max_lanes / 256));
    }
}

// Each grid covers the capacity; threads beyond their kind's range (read on the device) return at once. Calls with
// fewer than 2^17 points split each record's rows over 8 lanes (a choice made from the call's shape alone, so a
// call and its captured plan run the same split).
void launch_near_eval(hipStream_t s, const FieldParams &p)
{
    if (p.near_cap <= 0) return;
    if (p.n_pts < kNearSplitPoints) {
        launch_near_eval_k<false, kNearSplit>(s, p);
        if (p.float_records) launch_near_eval_k<true, kNearSplit>(s, p);
    } else {
        launch_near_eval_k<false, 1>(s, p);
        if (p.float_records) launch_near_eval_k<true, 1>(s, p);
    }
}

void launch_near_add(hipStream_t s, const FieldParams &p)
{
    if (p.n_pts <= 0) return;
    hipLaunchKernelGGL(near_add_kernel, dim3((unsigned)((p.n_pts + 255) / 256)), dim3(256), 0, s, p);
}

}  // namespace ccoil
