// Coaxial (z-axis) pair rules and self-inductance on the device: work items, parameters and launchers.
#pragma once

#include "common.hpp"

namespace ccoil {

enum : int32_t { kZMutual = 0, kZForce = 1, kZSelf = 2 };
constexpr int kZBlock = 256;   // threads per workgroup of zaxis_kernel

// One coaxial evaluation (a pair's M or axial force, or one coil's self-inductance). Every workgroup of the
// item covers kZBlock (radial node, angular node) units; each unit sums over all primary thickness nodes.
struct alignas(16) ZItemDev {
    double z[4];          // axial offsets constZ1..4 (mutual, force); z[0] = coil length (self)
    double zsq[4];        // their squares
    double r_inner;       // radial rule: r = r_inner + r_width (block + 0.5) + r_width / 2 x
    double r_width;
    double t_inner;       // primary thickness rule, same form
    double t_width;
    double a_width;       // angular block width pi / blocks
    double constant;      // per-node constant of the reference rule
    double mul, mul2, div;// result = ((sum * mul) * mul2) / div
    int32_t r_blk, r_inc, t_blk, t_inc, a_blk, a_inc;
    int32_t mode;         // kZMutual, kZForce, kZSelf
    int32_t out;          // index into ZParams::out
    int64_t first_wg;     // first workgroup of this item (ascending over items)
    int64_t units;        // radial nodes x angular nodes
};

struct ZParams {
    const ZItemDev *items;
    int32_t n_items;
    int64_t n_wg;
    double *partial;      // one per workgroup
    double *out;
    Tables tab;
};

void launch_zaxis(hipStream_t s, const ZParams &p);
void launch_zaxis_reduce(hipStream_t s, const ZParams &p);

}  // namespace ccoil
