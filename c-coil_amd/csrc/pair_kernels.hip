// Pair kernels: M, F and T of many (primary, secondary) pairs of mixed type and uneven cost in one launch.
//
// A work item is (pair, chunk of secondary nodes, group of the primary's thickness rows) and one wave (one
// 64-lane workgroup) owns one item: item = blockIdx.x. Items are laid out longest-first on the host, and
// workgroups are dispatched in index order, so the launch is an LPT schedule without a shared cursor. Each lane
// evaluates the primary's A or B (over the item's rows) at one secondary node at a time (the primary coil is
// wave-uniform, so its loops and tables are scalar), the weighted sums (M, F, T and the gradient of M, all linear
// in the row sums) are reduced across the wave in FP64 and one partial is written per item. A second kernel
// adds the partials of each pair in a fixed order, so results are deterministic. The reference instead runs one
// thread per (configuration, secondary node) and FP32 atomicAdds per configuration (MInductance.cu:157,
// Force.cu:172-178).
#include "integrands.hpp"
#include "kernels.hpp"
#include "multipole.hpp"
#include "pairs.hpp"

namespace ccoil {

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// PREC: 0 FP64 (stable integrands, refined hardware reciprocals), 1 FP32C (stable in FP32, near-winding
// nodes as in FP64), 2 FP32 (reference integrands in float), 3 FP64 reference integrands, 4 FP64 with IEEE
// division and square root.
template <int PREC, bool WM, bool WF, bool GR>
__device__ inline void eval_prim(const CoilDev &pc, const Tables &tab, double z, double r, int row0, int rstep,
                                 Local &L, AdjA *ada)
{
    if constexpr (GR) {   // gradients of M: A, B and the pair adjoint extras; FP64, FP64_IEEE or FP32C far nodes
        if constexpr (PREC == 4) eval_stable<double, true, true, false, false, true>(pc, tab, z, r, row0, rstep, L, nullptr, ada);
        else if constexpr (PREC == 1) eval_stable<float, true, true, false, false, true>(pc, tab, z, r, row0, rstep, L, nullptr, ada);
        else eval_stable<f64s, true, true, false, false, true>(pc, tab, z, r, row0, rstep, L, nullptr, ada);
        return;
    }
    if constexpr (PREC == 0) {
        eval_stable<f64s, WM, WF, false, false, false, true>(pc, tab, z, r, row0, rstep, L);
    } else if constexpr (PREC == 4) {
        eval_stable<double, WM, WF, false>(pc, tab, z, r, row0, rstep, L);
    } else if constexpr (PREC == 2) {
        eval_plain<float, WM, WF, false>(pc, tab, z, r, row0, rstep, L);
    } else if constexpr (PREC == 3) {
        eval_plain<double, WM, WF, false>(pc, tab, z, r, row0, rstep, L);
    } else if (is_near(pc, z, r, false)) {
        eval_stable<f64s, WM, WF, false>(pc, tab, z, r, row0, rstep, L);
    } else {
        eval_stable<float, WM, WF, false>(pc, tab, z, r, row0, rstep, L);
    }
}

__device__ inline void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Gradient of M with respect to the secondary's position: M = sum_n w A(x_n) . t_n over the secondary nodes x_n,
// so dM/dp = sum_n w J_A(x_n)^T t_n with J_A the Jacobian of the primary's A. With A = A_phi(r, z) e_phi, J_A follows
// from A_phi / r, dA_phi/dr (Ar: the rule's own radial derivative, summed node by node) and dA_phi/dz = -B_rho
// (exact node by node); on the axis A_phi / r and dA_phi/dr are equal.
__device__ inline void grad_m_term(const CoilDev &pc, const Local &L, double Ar, double r, double c, double s,
                                   const double *tan, double *out, double *tl_out = nullptr, double *gl_out = nullptr)
{
    double tl[3];
    mat_vec(pc.Tinv, tan, tl);
    double aor, dra;
    if (r / pc.R <= kGpuAxisRatio) {
        aor = Ar;
        dra = Ar;
    } else {
        aor = L.A / r;
        dra = Ar;
    }
    const double dd = dra - aor;
    const double g[3] = {tl[0] * (-c * s * dd) + tl[1] * (c * c * dra + s * s * aor),
                         tl[0] * (-s * s * dra - c * c * aor) + tl[1] * (c * s * dd),
                         (tl[0] * s - tl[1] * c) * L.Bh};
    mat_vec(pc.T, g, out);
    if (tl_out)
        for (int k = 0; k < 3; ++k) {
            tl_out[k] = tl[k];
            gl_out[k] = g[k];
        }
}

template <int PREC, bool WM, bool WF, bool WD, bool GR = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void pair_kernel(PairParams p)
{
    static_assert(!GR || (WM && !WF && !WD), "the gradient instance evaluates M and its derivatives");
    const int lane = threadIdx.x;
    const int64_t item = blockIdx.x;
    {
        int lo = 0, hi = p.n_pairs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (p.pairs[mid].first_item <= item) lo = mid;
            else hi = mid - 1;
        }
        const PairDev &pd = p.pairs[lo];
        const CoilDev &pc = p.prim[pd.prim];
        const CoilDev &sc = p.sec[pd.sec];
        const int64_t it_pair = item - pd.first_item;
        const int rg = (int)(it_pair % pd.row_groups);
        const int64_t n0 = (it_pair / pd.row_groups) * pd.nodes_per_item;
        const int64_t n1 = min(n0 + (int64_t)pd.nodes_per_item, pd.n_nodes);
        const int NA = sc.blk_ang * sc.n_ang, NT = sc.blk_thk * sc.n_thk;
        const bool table_ring = sc.blk_ang == 1 && pd.half_ring == 0 && pd.ring_offset == 0.0;
        const double ang_blk = (pd.half_ring ? kPi : 2.0 * kPi) / sc.blk_ang;
        const double tbs = sc.a / sc.blk_thk, lbs = sc.b / sc.blk_len;
        const double ax[3] = {sc.sa * sc.cb, sc.sa * sc.sb, sc.ca};
        const double u[3] = {sc.cb * sc.ca, sc.sb * sc.ca, -sc.sa};
        const double v[3] = {-sc.sb, sc.cb, 0.0};
        const int oa = table_offset(sc.n_ang), ot = table_offset(sc.n_thk), ol = table_offset(sc.n_len);
        double m = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
        double gs[kPairGradTerms - 1] = {};   // GR: the derivative sums after M
        // Secondary node k: weight, arm from the secondary's centre, unit tangent, the node in the primary's local
        // frame, and (GR) the ring angle's cos / sin, radius and length coordinate with their derivatives with
        // respect to the secondary's thickness and length.
        struct Node {
            double w, arm[3], tan[3], z, r, pcph, psph;
            double cs, ss, rho, ell, ht, hl;
        };
        auto node = [&](int64_t k, Node &n) {
            const int ia = (int)(k % NA);
            const int64_t rest = k / NA;
            const int it = (int)(rest % NT);
            const int il = (int)(rest / NT);
            const int i = ia % sc.n_ang, j = it % sc.n_thk, l = il % sc.n_len;
            double cph, sph;
            if (table_ring) {
                cph = p.tab.rcos[oa + i];
                sph = p.tab.rsin[oa + i];
            } else {
                const double phi = pd.ring_offset + ang_blk * ((ia / sc.n_ang) + 0.5) + 0.5 * ang_blk * p.tab.x[oa + i];
                sincos(phi, &sph, &cph);
            }
            const double rho = sc.R + tbs * ((it / sc.n_thk) + 0.5) + 0.5 * tbs * p.tab.x[ot + j];
            const double ell = -0.5 * sc.b + lbs * ((il / sc.n_len) + 0.5) + 0.5 * lbs * p.tab.x[ol + l];
            n.w = 0.125 * rho * p.tab.w[ol + l] * p.tab.w[ot + j] * p.tab.w[oa + i];
            if constexpr (GR) {
                n.cs = cph;
                n.ss = sph;
                n.rho = rho;
                n.ell = ell;
                n.ht = ((it / sc.n_thk) + 0.5 + 0.5 * p.tab.x[ot + j]) / sc.blk_thk;
                n.hl = -0.5 + ((il / sc.n_len) + 0.5 + 0.5 * p.tab.x[ol + l]) / sc.blk_len;
            }
            for (int q = 0; q < 3; ++q) {
                n.arm[q] = ell * ax[q] + rho * (cph * u[q] + sph * v[q]);
                n.tan[q] = -sph * u[q] + cph * v[q];
            }
            to_local(pc, sc.pos[0] + n.arm[0], sc.pos[1] + n.arm[1], sc.pos[2] + n.arm[2], n.z, n.r, n.pcph, n.psph);
        };
        // GR: the derivatives of M = sum_n w A(x_n) . t_n. With g = J_A^T t (the global Jacobian of the primary's A
        // at the node, as for dM/dp), the secondary's parameters move the node x and turn the tangent t:
        // d/dtheta (A . t) = (dx/dtheta) . g + A . dt/dtheta, and the weight w carries a factor rho. The ring frame
        // (ax, u, v) turns under alpha about v (d ax = u, d u = -ax) and under beta about the z axis (d ax = sa v,
        // d u = ca v, d v = -(ca u + sa ax)). The primary's angles turn its field: d A = T (Omega A_l - J_l Omega x_l)
        // with Omega x = om x x; its R, a, b change A_phi by the rows' homogeneity (AdjA).
        auto accumulate_grad = [&](const Node &n, const Local &L, const AdjA &q) {
            double gd[3], tl[3], gl[3];
            grad_m_term(pc, L, q.Ar, n.r, n.pcph, n.psph, n.tan, gd, tl, gl);
            const double Al[3] = {-L.A * n.psph, L.A * n.pcph, 0.0};
            double Ag[3];
            mat_vec(pc.T, Al, Ag);
            const double At = dot3(Ag, n.tan);
            m += n.w * At;
            gs[0] += n.w * gd[0];
            gs[1] += n.w * gd[1];
            gs[2] += n.w * gd[2];
            const double er[3] = {n.cs * u[0] + n.ss * v[0], n.cs * u[1] + n.ss * v[1], n.cs * u[2] + n.ss * v[2]};
            const double rad = At / n.rho + dot3(er, gd);            // d(w A . t)/d rho / w
            double dxa[3], dta[3], dxb[3], dtb[3];
            const double cav[3] = {sc.ca * u[0] + sc.sa * ax[0], sc.ca * u[1] + sc.sa * ax[1], sc.ca * u[2] + sc.sa * ax[2]};
            for (int q3 = 0; q3 < 3; ++q3) {
                dxa[q3] = n.ell * u[q3] - n.rho * n.cs * ax[q3];
                dta[q3] = n.ss * ax[q3];
                dxb[q3] = n.ell * sc.sa * v[q3] + n.rho * (n.cs * sc.ca * v[q3] - n.ss * cav[q3]);
                dtb[q3] = -n.ss * sc.ca * v[q3] - n.cs * cav[q3];
            }
            gs[3] += n.w * (dot3(dxa, gd) + dot3(Ag, dta));
            gs[4] += n.w * (dot3(dxb, gd) + dot3(Ag, dtb));
            gs[5] += n.w * rad;
            gs[6] += n.w * n.ht * rad;
            gs[7] += n.w * n.hl * dot3(ax, gd);
            // primary angles: t_l . (om x A_l) - (om x x_l) . g_l, x_l the node in the primary's frame
            const double xl[3] = {n.r * n.pcph, n.r * n.psph, n.z};
            double oa[3], ob[3];
            cross3(pc.om_a, Al, oa);
            cross3(pc.om_b, Al, ob);
            double xa[3], xb[3];
            cross3(pc.om_a, xl, xa);
            cross3(pc.om_b, xl, xb);
            gs[8] += n.w * (dot3(tl, oa) - dot3(xa, gl));
            gs[9] += n.w * (dot3(tl, ob) - dot3(xb, gl));
            // primary geometry: A_phi's derivatives along e_phi
            const double tphi = -tl[0] * n.psph + tl[1] * n.pcph;
            gs[10] += n.w * tphi * q.dR;
            gs[11] += n.w * tphi * q.da;
            gs[12] += n.w * tphi * q.db;
        };
        // Adds the node's weighted terms: M from A, F and T from B, the gradient of M from both.
        auto accumulate = [&](const Node &n, const Local &L, const AdjA &q) {
            if (WM) {
                const double vl[3] = {-L.A * n.psph, L.A * n.pcph, 0.0};
                double ag[3];
                mat_vec(pc.T, vl, ag);
                m += n.w * (ag[0] * n.tan[0] + ag[1] * n.tan[1] + ag[2] * n.tan[2]);
            }
            if (WF) {
                const double vl[3] = {L.Bh * n.pcph, L.Bh * n.psph, L.Bz};
                double bg[3];
                mat_vec(pc.T, vl, bg);
                const double df0 = n.w * (n.tan[1] * bg[2] - n.tan[2] * bg[1]);
                const double df1 = n.w * (n.tan[2] * bg[0] - n.tan[0] * bg[2]);
                const double df2 = n.w * (n.tan[0] * bg[1] - n.tan[1] * bg[0]);
                f0 += df0;
                f1 += df1;
                f2 += df2;
                t0 += n.arm[1] * df2 - n.arm[2] * df1;
                t1 += n.arm[2] * df0 - n.arm[0] * df2;
                t2 += n.arm[0] * df1 - n.arm[1] * df0;
            }
            if (WD) {
                double gd[3];
                grad_m_term(pc, L, q.Ar, n.r, n.pcph, n.psph, n.tan, gd);
                d0 += n.w * gd[0];
                d1 += n.w * gd[1];
                d2 += n.w * gd[2];
            }
        };
        // Secondary nodes beyond the primary's far-field distance take the primary's series (FP64): once per node,
        // in row group 0 (the other groups of a pair skip them). The loop runs to the wave's largest term count and
        // each lane stops adding at its own, so a node's value does not depend on its wave. Not in the gradient
        // instance (the series carries no derivatives with respect to the primary's R, a, b).
        auto series = [&](const Node &n, bool far) {
            if constexpr (!GR) {
                if (!__ballot(far)) return;
                const int own = far ? multipole_terms(n.z * n.z + n.r * n.r, pc.inv_rmax2,
                                                      PREC == 1 ? kMpLogTolFP32C : kMpLogTolFP64) : 0;
                const int terms = wave_max_small(own), lo = wave_min_small(own, far);
                if (far && rg == 0) {
                    Local L = {0, 0, 0, 0, 0, 0, 0};
                    double aor = 0.0;
                    eval_multipole_pair<WM || WD, WF || WD>(p.mp + kMpHeader + (size_t)pd.prim * kMpCoil, p.mp,
                                                            pc.rmax, n.z, n.r, terms, lo, own, L, aor);
                    AdjA q = {0.0, 0.0, 0.0, L.Bz - aor};   // dA_phi/dr = B_z - A_phi / r
                    accumulate(n, L, q);
                }
            }
        };
        if constexpr (PREC == 1) {
            // FP32C: far nodes in FP32 as they come; nodes in the primary's near-winding shell are queued (ring of
            // 128 in LDS, wave-uniform head and count) and evaluated in FP64 64 at a time, so the FP64 evaluation
            // runs with full waves instead of beside the FP32 one in every wave that holds a near node.
            __shared__ int64_t queue[128];
            int qh = 0, qn = 0;
            for (int64_t base = n0; base < n1; base += 64) {
                const int64_t k = base + lane;
                Node n;
                bool near = false, far = false;
                if (k < n1) {
                    node(k, n);
                    far = !GR && p.mp && is_far(pc, n.z, n.r);
                    near = !far && is_near(pc, n.z, n.r, WD || GR);   // the position gradient divides A by r
                    if (!near && !far) {
                        Local L = {0, 0, 0, 0, 0, 0, 0};
                        if constexpr (GR) {
                            AdjA q;
                            eval_stable<float, true, true, false, false, true>(pc, p.tab, n.z, n.r, rg, pd.row_groups,
                                                                               L, nullptr, &q);
                            accumulate_grad(n, L, q);
                        } else {
                            AdjA q;
                            eval_stable<float, WM || WD, WF || WD, false, false, WD, true>(pc, p.tab, n.z, n.r, rg,
                                                                                     pd.row_groups, L, nullptr, &q);
                            accumulate(n, L, q);
                        }
                    }
                }
                series(n, far);
                const unsigned long long mask = __ballot(near);
                if (mask) {
                    if (near) queue[(qh + qn + __popcll(mask & ((1ull << lane) - 1ull))) & 127] = k;
                    qn += __popcll(mask);
                    __syncthreads();
                    if (qn >= 64) {
                        node(queue[(qh + lane) & 127], n);
                        Local L = {0, 0, 0, 0, 0, 0, 0};
                        if constexpr (GR) {
                            AdjA q;
                            eval_stable<f64s, true, true, false, false, true>(pc, p.tab, n.z, n.r, rg, pd.row_groups,
                                                                              L, nullptr, &q);
                            accumulate_grad(n, L, q);
                        } else {
                            AdjA q;
                            eval_stable<f64s, WM || WD, WF || WD, false, false, WD, true>(pc, p.tab, n.z, n.r, rg,
                                                                                    pd.row_groups, L, nullptr, &q);
                            accumulate(n, L, q);
                        }
                        qh = (qh + 64) & 127;
                        qn -= 64;
                    }
                    __syncthreads();
                }
            }
            if (qn > 0 && lane < qn) {
                Node n;
                node(queue[(qh + lane) & 127], n);
                Local L = {0, 0, 0, 0, 0, 0, 0};
                if constexpr (GR) {
                    AdjA q;
                    eval_stable<f64s, true, true, false, false, true>(pc, p.tab, n.z, n.r, rg, pd.row_groups, L,
                                                                      nullptr, &q);
                    accumulate_grad(n, L, q);
                } else {
                    AdjA q;
                    eval_stable<f64s, WM || WD, WF || WD, false, false, WD, true>(pc, p.tab, n.z, n.r, rg,
                                                                                  pd.row_groups, L, nullptr, &q);
                    accumulate(n, L, q);
                }
            }
        } else {
            for (int64_t k = n0 + lane; k < n1; k += 64) {
                Node n;
                node(k, n);
                const bool far = !GR && p.mp && is_far(pc, n.z, n.r);
                if (!far) {
                    Local L = {0, 0, 0, 0, 0, 0, 0};
                    AdjA q;
                    eval_prim<PREC, WM || WD, WF || WD, GR || WD>(pc, p.tab, n.z, n.r, rg, pd.row_groups, L, &q);
                    if constexpr (GR) accumulate_grad(n, L, q);
                    else accumulate(n, L, q);
                }
                series(n, far);
            }
        }
        if constexpr (GR) {
            double *out = p.partial + kPairGradTerms * item;
            m = wave_sum(m);
            if (lane == 0) out[0] = m;
            for (int q = 0; q < kPairGradTerms - 1; ++q) {
                const double v = wave_sum(gs[q]);
                if (lane == 0) out[1 + q] = v;
            }
            return;
        }
        double *out = p.partial + kPairTerms * item;
        if (WM) {
            m = wave_sum(m);
            if (lane == 0) out[0] = m;
        }
        if (WF) {
            f0 = wave_sum(f0); f1 = wave_sum(f1); f2 = wave_sum(f2);
            t0 = wave_sum(t0); t1 = wave_sum(t1); t2 = wave_sum(t2);
            if (lane == 0) {
                out[1] = f0; out[2] = f1; out[3] = f2;
                out[4] = t0; out[5] = t1; out[6] = t2;
            }
        }
        if (WD) {
            d0 = wave_sum(d0); d1 = wave_sum(d1); d2 = wave_sum(d2);
            if (lane == 0) {
                out[7] = d0; out[8] = d1; out[9] = d2;
            }
        }
    }
}

// The gradient of M: one wave per pair, lane l adds items l, l + 64, ..., a fixed butterfly combines the lanes.
__global__ __launch_bounds__(64) void pair_grad_reduce_kernel(PairParams p)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    const PairDev &pd = p.pairs[k];
    const int64_t end = (k + 1 < p.n_pairs) ? p.pairs[k + 1].first_item : p.n_items;
    double s[kPairGradTerms];
    for (int q = 0; q < kPairGradTerms; ++q) s[q] = 0.0;
    for (int64_t it = pd.first_item + lane; it < end; it += 64)
        for (int q = 0; q < kPairGradTerms; ++q) s[q] += p.partial[kPairGradTerms * it + q];
    for (int q = 0; q < kPairGradTerms; ++q) s[q] = wave_sum(s[q]);
    if (lane != 0) return;
    const double ms = pd.m_scale;
    if (p.M) p.M[pd.out] = s[0] * ms;
    double *g = p.grad + (int64_t)kPairGradOut * pd.out;
    for (int q = 0; q < 8; ++q) g[q] = s[1 + q] * ms;                        // secondary x y z alpha beta R a b
    for (int q = 0; q < 3; ++q) g[8 + q] = -s[1 + q] * ms;                   // primary position
    for (int q = 0; q < 5; ++q) g[11 + q] = s[9 + q] * ms;                   // primary alpha beta R a b
}

__global__ __launch_bounds__(64) void pair_reduce_kernel(PairParams p)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    const PairDev &pd = p.pairs[k];
    const int64_t end = (k + 1 < p.n_pairs) ? p.pairs[k + 1].first_item : p.n_items;
    double s[kPairTerms];
    for (int q = 0; q < kPairTerms; ++q) s[q] = 0.0;
    for (int64_t it = pd.first_item + lane; it < end; it += 64)
        for (int q = 0; q < kPairTerms; ++q) s[q] += p.partial[kPairTerms * it + q];
    for (int q = 0; q < kPairTerms; ++q) s[q] = wave_sum(s[q]);
    if (lane != 0) return;
    if (p.what & 1) p.M[pd.out] = s[0] * pd.m_scale;
    if (p.what & 2) {
        for (int q = 0; q < 3; ++q) p.F[3 * pd.out + q] = s[1 + q] * pd.f_scale;
        for (int q = 0; q < 3; ++q) p.T[3 * pd.out + q] = s[4 + q] * pd.f_scale;
    }
    if (p.what & 4)
        for (int q = 0; q < 3; ++q) p.dM[3 * pd.out + q] = s[7 + q] * pd.m_scale;
}

template <int PREC>
static void launch_pairs_p(hipStream_t s, const PairParams &p)
{
    const bool wm = p.what & 1, wf = p.what & 2, wd = p.what & 4;
    const dim3 grid((unsigned)p.n_items);
    if (p.what & kPairGrad) {
        hipLaunchKernelGGL((pair_kernel<PREC, true, false, false, true>), grid, dim3(64), 0, s, p);
    } else if (wd) {
        // the gradient needs A and B at every node; M, F and T come with it at little extra cost
        hipLaunchKernelGGL((pair_kernel<PREC, true, true, true>), grid, dim3(64), 0, s, p);
    } else if (wm && wf) {
        hipLaunchKernelGGL((pair_kernel<PREC, true, true, false>), grid, dim3(64), 0, s, p);
    } else if (wm) {
        hipLaunchKernelGGL((pair_kernel<PREC, true, false, false>), grid, dim3(64), 0, s, p);
    } else {
        hipLaunchKernelGGL((pair_kernel<PREC, false, true, false>), grid, dim3(64), 0, s, p);
    }
}

void launch_pairs(hipStream_t s, const PairParams &p)
{
    if (p.n_items <= 0) return;
    if (p.mp) launch_multipole_moments(s, p.prim, p.n_prim, p.tab, p.mp);
    if (p.precision == 0) launch_pairs_p<0>(s, p);
    else if (p.precision == 4) launch_pairs_p<4>(s, p);
    else if (p.precision == 2) launch_pairs_p<2>(s, p);
    else if (p.precision == 3) launch_pairs_p<3>(s, p);
    else launch_pairs_p<1>(s, p);
}

void launch_pair_reduce(hipStream_t s, const PairParams &p)
{
    if (p.n_pairs <= 0) return;
    if (p.what & kPairGrad) hipLaunchKernelGGL(pair_grad_reduce_kernel, dim3((unsigned)p.n_pairs), dim3(64), 0, s, p);
    else hipLaunchKernelGGL(pair_reduce_kernel, dim3((unsigned)p.n_pairs), dim3(64), 0, s, p);
}

}  // namespace ccoil
