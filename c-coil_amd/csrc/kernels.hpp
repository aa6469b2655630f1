// Kernel parameter blocks and launcher declarations (host <-> device contract inside the library).
#pragma once

#include "common.hpp"

namespace ccoil {

// Arithmetic of the quadrature (values of the C ABI precision argument).
enum { kFP64 = 0, kFP32C = 1, kFP32 = 2, kFP64Reference = 3, kFP64IEEE = 4 };
constexpr int kFieldBlock = 256;

// Device-side outcome of the FP32C near pass (read back by the host once the call's work is done).
struct NearState {
    unsigned long long total;   // records (deferred (coil, point) pairs) counted by field_kernel
    unsigned int overflow;      // total > near_cap: the pass was skipped
    unsigned int mismatch;      // near_fill classified a pair differently from field_kernel
    unsigned long long fp64;    // records of kind 0 (FP64 rule)
};

struct FieldParams {
    const CoilDev *coils;
    int32_t n_coils;
    const double *pts;     // AoS xyz
    int64_t n_pts;
    double *A, *E, *B, *G; // AoS outputs (NULL = not requested)
    uint32_t q;            // Q_* mask
    Tables tab;
    int32_t near_fp64;     // FP32C: skip near pairs here, re-evaluate in FP64
    // Rule records (the series modes, large calls): field_kernel evaluates only the series and defers every other
    // (coil, point) pair to the record pass below, which runs the rule on waves of one coil each. A record's kind is
    // 0 for the FP64 rule (FP32C's near pairs; every rule pair in FP64) and 1 for the float rule (FP32C's others);
    // its key is kind * n_coils + CoilDev::near_key, n_keys = n_coils * (1 + float_records).
    int32_t rule_records;
    int32_t float_records;
    int32_t n_keys;
    // FP32C far field: the recurrence header and each coil's moments (kMpHeader + n_coils * kMpCoil doubles,
    // multipole.hpp); a coil with CoilDev::far_d2 > 0 takes its series beyond that distance. NULL: no coil does.
    const double *mp;
    // The order the points are visited in (a spatial sort, so the points of a wave lie close together and a wave
    // mostly takes one branch per coil); NULL: index order. No result depends on it.
    const uint32_t *order;
    // FP32C near-winding pass, stream-ordered with no host readback (so a caller may capture it in a graph):
    // field_kernel counts each point's near coils (near_counts); an exclusive scan of near_counts gives each point's
    // slots (near_offsets); near_count counts each coil's near points (near_hist, indexed by CoilDev::near_key, so
    // the costliest coils come first); near_key_scan turns near_hist into per-coil cursors, and stores the total and whether
    // it exceeds near_cap in near_state; near_fill places one record per (coil, point) pair at its coil's cursor:
    // the coil key (near_keys) and point << 32 | slot (near_recs); near_eval writes each record's FP64
    // contribution into its slot of near_vals (near_width doubles: the requested A, E, B, G in that order);
    // near_add adds a point's slots in slot (coil) order. The order of records within a coil varies from run to
    // run, but a slot's value depends only on its (coil, point) and every sum has a fixed order, so FP32C results
    // are bitwise reproducible. When the total exceeds near_cap, near_fill, near_eval and near_add do nothing and
    // the host grows the buffers and runs them again.
    int64_t *near_counts;
    const int64_t *near_offsets;
    uint32_t *near_hist;      // [n_keys + 1] counts per key
    uint32_t *near_cursor;    // [n_keys + 1] next free record of each key
    NearState *near_state;
    int64_t near_cap;
    uint32_t *near_keys;
    const uint32_t *near_coil;               // near_key -> coil index
    uint64_t *near_recs;
    double *near_vals;
    int32_t near_width;
    // Coil chunks (grid y): each block row sums its chunk of coils into partial ([chunk][point][width]: the
    // requested outputs A, E, B, G in that order, 3 or 9 doubles each) and field_reduce_kernel adds the chunks in
    // order.
    int32_t coils_per_chunk;
    int32_t n_chunks;
    int32_t partial_width;
    double *partial;
};

// What becomes of a (coil, point) pair: -1 field_kernel evaluates it (the series beyond the far-field distance, or
// the rule in the kernel), 0 an FP64-rule record, 1 a float-rule record. field_kernel, near_count and near_fill all
// classify with this, bitwise alike (to_local and the tests round operation by operation).
__device__ inline int record_kind(const FieldParams &p, const CoilDev &c, double z, double r, bool axis)
{
    if (is_far(c, z, r)) return -1;
    const bool near = p.near_fp64 && is_near(c, z, r, axis);
    if (p.rule_records) return (p.float_records && !near) ? 1 : 0;
    return near ? 0 : -1;
}

// Adjoint of the summed B field: per-coil gradients of sum_x dL/dB(x) . B(x), 9 per coil in the order
// current, R, a, b, px, py, pz, alpha, beta. Grid: (point blocks, coils); partial sums per (coil, block).
constexpr int kGradTerms = 9;
struct GradParams {
    const CoilDev *coils;
    int32_t n_coils;
    const double *pts;     // AoS xyz
    int64_t n_pts;
    const double *gB;      // dL/dB, AoS (global frame)
    double *partial;       // [n_coils][n_blocks][kGradTerms]
    int32_t n_blocks;
    double *grad;          // [n_coils][kGradTerms]
    Tables tab;
};
void launch_field_grad(hipStream_t s, const GradParams &p, int precision);

void launch_gl_tables(hipStream_t s, double *x, double *w, double *cosv, double *omc, float *xf, float *wf,
                      float *cosf, float *omcf, double *rcos, double *rsin);
// Per-entry products of the tables (after launch_gl_tables on the same stream): wc = w cos, opc = 1 + cos,
// ws = w (1 - cos), ws2 = w (1 - cos)^2, wcs = w cos (1 - cos) (and its float copy wcsf).
void launch_angular_tables(hipStream_t s, const double *x, const double *w, const double *cosv, const double *omc,
                           double *wc, float *wcf, double *opc, float *opcf, double *ws, double *ws2, double *wcs,
                           float *wcsf);
// split: 1, 4 or 16 lanes per point.
// Zeroes n_words 8-byte words with a kernel (no hipMemsetAsync, so callers may capture it in a graph).
void launch_fill_zero(hipStream_t s, void *p, int64_t n_words);
void launch_field(hipStream_t s, const FieldParams &p, int precision, int split);
void launch_field_reduce(hipStream_t s, const FieldParams &p);   // after launch_field when n_chunks > 1
// Resident 256-thread blocks per CU of the field kernel instance (0 if the runtime cannot tell).
int field_blocks_per_cu(uint32_t q, int precision, int split);
// Whether the series-only kernel counts the record pass's keys (in LDS, 4 bytes per coil of a chunk and kind), so
// that near_count_kernel is not launched: calls with rule records whose coil chunks fit.
__host__ __device__ inline bool series_counts_keys(const FieldParams &p)
{
    return p.rule_records && p.coils_per_chunk <= 4096;
}
// Resident 256-thread blocks per CU of the series-only kernel (the launch of calls with rule records).
int field_series_blocks_per_cu(uint32_t q, int precision);
// Exclusive prefix sum of n counts into offsets (kernel launches only); with temp == NULL only sets temp_bytes.
hipError_t near_scan(hipStream_t s, const int64_t *counts, int64_t *offsets, int64_t n, void *temp,
                     size_t &temp_bytes);
// near_count: per-coil counts into near_hist (zeroed before); near_key_scan: per-coil cursors from near_hist, and
// near_state's total and overflow (one small kernel).
void launch_near_count(hipStream_t s, const FieldParams &p);
void launch_near_key_scan(hipStream_t s, const FieldParams &p);
void launch_near_fill(hipStream_t s, const FieldParams &p);
void launch_near_eval(hipStream_t s, const FieldParams &p);
void launch_near_add(hipStream_t s, const FieldParams &p);
// Far-field moments of every coil with CoilDev::far_d2 > 0 and the recurrence header, into mp.
void launch_multipole_moments(hipStream_t s, const CoilDev *coils, int n_coils, const Tables &tab, double *mp);
// The visit order: points counting-sorted by the Morton cell of a 2^bits-per-axis grid over their bounding box.
// temp: visit_order_bytes(n, bits, scan_bytes) bytes, scan_bytes from near_scan for 2^(3 bits) counts.
size_t visit_order_bytes(int64_t n, int bits, size_t scan_bytes);
hipError_t launch_visit_order(hipStream_t s, const double *pts, int64_t n, int bits, void *temp, size_t scan_bytes,
                              uint32_t *order);

}  // namespace ccoil
