// Pair (mutual inductance, force, torque) work queue: parameters, buffers and launchers.
#pragma once

#include "common.hpp"

namespace ccoil {

// Per-pair description, sorted by cost (longest first) on the host.
struct PairDev {
    int32_t prim;          // index into the primary coil array
    int32_t sec;           // index into the secondary coil array
    int32_t out;           // output slot (original pair index)
    int32_t nodes_per_item;
    int64_t first_item;    // first work item of this pair
    int64_t n_nodes;       // secondary nodes of this pair
    double ring_offset;    // improved-precision angular offset (SupportFunctions.cxx:152-157)
    int32_t half_ring;     // 1: integrate phi over [offset, offset + pi] (SupportFunctions.cxx:86-87)
    int32_t row_groups;    // the primary's thickness rows are split over this many items (rows g, g + G, ...)
    double m_scale;        // 2 pi N2 / (I1 * blocks)   (CalculateMInductanceGeneral.cxx:115-116)
    double f_scale;        // 2 pi N2 I2 / blocks, or 2 pi N2 / I1 (reference arrangement scale)
};

struct PairParams {
    const CoilDev *prim;
    const CoilDev *sec;
    const PairDev *pairs;
    int32_t n_pairs;
    int64_t n_items;
    double *partial;                 // kPairTerms doubles per item: M, F[3], T[3], dM/dp[3]
    uint32_t what;                   // 1: M, 2: F and T, 4: dM/d(secondary position), 8: the full gradient of M
    int32_t precision;
    Tables tab;
    double *M, *F, *T, *dM;          // outputs (device), indexed by PairDev::out
    double *grad;                    // what & 8: kPairGradOut doubles per pair
    // The primaries' far-field series (multipole.hpp; moments computed by launch_pairs before the pair kernel):
    // a secondary node beyond the primary's far-field distance takes the series instead of the rule. NULL: none.
    double *mp;
    int32_t n_prim;
};
constexpr int kPairTerms = 10;
// The gradient of M (what & 8): per item M and 13 derivative sums, dM/d(secondary x, y, z, alpha, beta, R, a, b,
// primary alpha, beta, R, a, b); per pair 16 outputs, the secondary's 8 then the primary's 8 (x, y, z, alpha, beta,
// R, a, b), the primary's position derivative being minus the secondary's.
constexpr int kPairGradTerms = 14;
constexpr int kPairGradOut = 16;
constexpr uint32_t kPairGrad = 8;

// One 64-lane workgroup per work item (grid = n_items, at most 2^31 - 1). A pair's items are its chunks of
// secondary nodes times its primary row groups: item i of a pair takes chunk i / row_groups and row group
// i % row_groups.
void launch_pairs(hipStream_t s, const PairParams &p);
// One 64-lane workgroup per pair: lane l adds items l, l + 64, ... and a fixed butterfly combines the lanes.
void launch_pair_reduce(hipStream_t s, const PairParams &p);

}  // namespace ccoil
