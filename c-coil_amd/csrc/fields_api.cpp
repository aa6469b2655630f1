// C ABI: ccoil_fields (include/ccoil_hip.h) and the body it shares with the other field entry points: validation,
// the coil records, the launch shape, the buffers (grow-only, or parts of a fixed arena), the kernels and the FP32C
// near pass.
#include "context.hpp"
#include "multipole.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

using namespace ccoil;

namespace ccoil {

// Launch-shape rule of a field call. Split the thickness rows over 4 lanes when there are too few points to give
// every SIMD of the 256 CUs several waves, and over 16 lanes when even one chunk per coil would leave the launch
// below one wave per SIMD (few points and few coils: then each lane's serial chain of nodes, not the issue rate,
// sets the kernel's duration; one coil at 4,096 points measured 46 us at 4 lanes and 25 us at 16, while at 65,536
// points 16 lanes were slower). Split the coils into chunks (grid y, summed in order by field_reduce_kernel) while
// the launch still has fewer than about 8 waves per SIMD. Large launches take 1-4 chunks, whichever makes the last
// round of resident waves fullest: every wave does the same work, so with R rounds of resident waves the launch
// runs ceil(R) rounds.
FieldShape field_shape_rule(ccoil_ctx ctx, int64_t n, int64_t n_coils, uint32_t q, int precision, bool series_only)
{
    FieldShape sh;
    const int64_t target = (int64_t)256 * 4 * 8;
    sh.split = n < (int64_t)256 * 4 * 64 * 4 ? 4 : 1;
    if (sh.split == 4 && (n * 4 + 63) / 64 * n_coils < (int64_t)256 * 4) sh.split = 16;
    if (series_only) sh.split = 1;   // no rows to split
    const int64_t waves = std::max<int64_t>(1, (n * sh.split + 63) / 64);
    int64_t chunks = waves >= target ? 1 : std::min<int64_t>(n_coils, (target + waves - 1) / waves);
    if (waves >= target && n * 18 < ((int64_t)1 << 31)) {   // the reduction grid stays below 2^31 threads
        const int bpc = series_only ? field_series_blocks_per_cu(q, precision)
                                    : field_blocks_per_cu(q, precision, sh.split);
        const double slots = (double)ctx->cu_count * bpc * (kFieldBlock / 64);
        if (slots > 0) {
            double best = 0.0;
            for (int64_t c = 1; c <= std::min<int64_t>(4, n_coils); ++c) {
                const double R = (double)waves * c / slots, eff = R / std::ceil(R);
                if (eff > best + 0.02) {
                    best = eff;
                    chunks = c;
                }
            }
        }
    }
    sh.chunks = (int)std::max<int64_t>(1, std::min<int64_t>(chunks, 65535));
    return sh;
}

// Distance, in units of the coil's r_max, beyond which the far-field series replaces the rule (0: never), at least
// ylo (far_distance below).
//
// The series sums the rule's own rings exactly, so it departs from the rule by the angular rule's error and its
// truncation (below 1e-15 from D = 1.17 r_max on). The angular integrand over phi in [0, pi] (Gauss-Legendre on
// phi = pi (1 + x) / 2, n_ang nodes) is singular where the distance from the point to the ring vanishes, at
// phi = +-i eta with cosh(eta) = 1 + d^2 / (2 r rho), d the point's distance from the ring in the meridian plane;
// for every point at distance D from the centre and every ring of the winding, d >= D - r_max and r rho <= D (R + a),
// so eta >= acosh(1 + (D - r_max)^2 / (2 D (R + a))). The singularity sits at x_s = -1 + 2 i eta / pi, outside the
// Bernstein ellipse of parameter rho_B = |x_s - sqrt(x_s^2 - 1)| (the root beyond 1), and the rule's error falls like
// rho_B^(-2 n_ang). Relative to the coil's field at D the error is at most 1000 y^2 rho_B^(-2 n_ang) (y = D / r_max;
// the y^2 for the node terms that cancel down to the far field, the constant 10x above what a filament, a thin and a
// thick coil measured in FP64 against the series at n_ang = 4..24 and y = 2..50). The threshold is the least y in
// [ylo, 100] with that bound below tol: for n_ang >= 5 the bound falls monotonically there, so g(y) = ln(bound / tol)
// has one sign change, found by regula falsi with the Illinois modification (a few evaluations where a bisection
// takes forty; a pair call sets one distance per primary) to within 1e-12 of y, returning the side where the bound
// holds.
double far_threshold(int n_ang, double q, double tol, double ylo)
{
    auto g = [&](double y) {
        const double d = y - 1.0;
        const double eta = std::acosh(1.0 + d * d * q / (2.0 * y));
        const std::complex<double> xs(-1.0, 2.0 * eta / kPi);
        const std::complex<double> sq = std::sqrt(xs * xs - 1.0);
        const double rho = std::max(std::abs(xs + sq), std::abs(xs - sq));
        return std::log(1000.0 / tol) + 2.0 * std::log(y) - 2.0 * n_ang * std::log(rho);
    };
    if (n_ang < 5) return 0.0;
    double lo = ylo, hi = 100.0, glo = g(lo), ghi = g(hi);
    if (glo <= 0.0) return ylo;
    if (ghi > 0.0) return 0.0;
    int side = 0;
    for (int it = 0; it < 200 && hi - lo > 1e-12 * hi; ++it) {
        double y = (lo * ghi - hi * glo) / (ghi - glo);
        if (!(y > lo && y < hi)) y = 0.5 * (lo + hi);
        const double gy = g(y);
        if (gy <= 0.0) {
            hi = y;
            ghi = gy;
            if (side == -1) glo *= 0.5;
            side = -1;
        } else {
            lo = y;
            glo = gy;
            if (side == 1) ghi *= 0.5;
            side = 1;
        }
    }
    return hi;
}

// Nearest distance, in units of r_max, at which the series may take over on cost: a pair's series of L(y) terms
// (multipole_terms: L = 2 ln(2 / tol) / ln(y^2) + 2) costs about 15 VALU operations per term plus 200 in the kernels
// of this library, the coil's rule about 55 per fast-method node and 12 per slow-method node (n_ang x n_thk nodes).
// Where the rule is dearer the series starts at the nearest distance with L(y) within budget, but not below kFarMinY
// (where L reaches about kMpTerms) and not beyond 2 r_max, from where the series replaces every rule.
double far_cost_floor(int n_ang, int n_thk, bool fast, double log_tol)
{
    const double rule = (double)n_ang * n_thk * (fast ? 55.0 : 12.0);
    if (rule <= 230.0) return 2.0;
    const double y = std::exp(15.0 * log_tol / (rule - 230.0));
    return std::min(2.0, std::max(kFarMinY, y));
}

int far_terms(double y, bool fp64)
{
    if (!(y > 0.0)) return 0;
    // multipole_terms at D = y r_max (it falls with D), in the same float arithmetic, one above for its fast log
    const float lg = std::log((float)(y * y));
    const int L = (int)(2.0f * (fp64 ? kMpLogTolFP64 : kMpLogTolFP32C) / lg) + 3;
    return std::min(kMpTerms, std::max(4, L));
}

double far_distance(const CoilDev &c, bool fp64, bool cost_floor)
{
    const double log_tol = fp64 ? kMpLogTolFP64 : kMpLogTolFP32C;
    return far_threshold(c.n_ang, c.rmax / (c.R + c.a), fp64 ? kFarToleranceFP64 : kFarTolerance,
                         cost_floor ? far_cost_floor(c.n_ang, c.n_thk, c.use_fast != 0, log_tol) : 2.0);
}

bool field_record_pass(int32_t precision, uint32_t flags, int64_t n_points, bool fixed_arena, bool *rule_records)
{
    const bool far = (precision == CCOIL_FP32C && !(flags & CCOIL_DIRECT)) ||
                     (precision == CCOIL_FP64 && (flags & CCOIL_FAR_FIELD));
    const bool rr = far && !fixed_arena && n_points >= kOrderMinPoints && n_points <= kRecordMaxPoints;
    if (rule_records) *rule_records = rr;
    return rr || (precision == CCOIL_FP32C && !(flags & CCOIL_NO_NEAR_FP64));
}

// Sizes the near pass's buffers of L for cap records and points L.p at them: the context's grow-only buffers, or
// parts of arena (then cap is lowered to what the arena still holds).
ccoil_status near_buffers(ccoil_ctx ctx, FieldLaunch &L, int64_t cap, int width, Arena *arena)
{
    FieldParams &p = L.p;
    const int64_t n = p.n_pts;
    void *counts, *offsets, *hist, *state, *temp, *keys, *recs, *vals;
    size_t tb = 0;
    CC_HIP(ctx, near_scan(ctx->stream, nullptr, nullptr, n, nullptr, tb));
    ccoil_status st;
    if ((st = call_buffer(ctx, ctx->ncounts, arena, sizeof(int64_t) * (size_t)n, &counts)) != CCOIL_OK) return st;
    if ((st = call_buffer(ctx, ctx->noffsets, arena, sizeof(int64_t) * (size_t)n, &offsets)) != CCOIL_OK) return st;
    if ((st = call_buffer(ctx, ctx->nhist, arena, 2 * sizeof(uint32_t) * ((size_t)p.n_keys + 2), &hist)) != CCOIL_OK)
        return st;
    if ((st = call_buffer(ctx, ctx->nstate, arena, sizeof(NearState), &state)) != CCOIL_OK) return st;
    if ((st = call_buffer(ctx, ctx->ntemp, arena, tb > 0 ? tb : 1, &temp)) != CCOIL_OK) return st;
    const size_t per_record = sizeof(uint32_t) + sizeof(uint64_t) + sizeof(double) * (size_t)width;
    if (arena) cap = std::min<int64_t>(cap, (int64_t)((arena->left() > 768 ? arena->left() - 768 : 0) / per_record));
    cap = std::max<int64_t>(cap, 1);
    if ((st = call_buffer(ctx, ctx->nkeys, arena, sizeof(uint32_t) * (size_t)cap, &keys)) != CCOIL_OK) return st;
    if ((st = call_buffer(ctx, ctx->nrecs, arena, sizeof(uint64_t) * (size_t)cap, &recs)) != CCOIL_OK) return st;
    if ((st = call_buffer(ctx, ctx->nvals, arena, sizeof(double) * (size_t)width * (size_t)cap, &vals)) != CCOIL_OK)
        return st;
    if (!arena) ctx->near_cap = std::max(ctx->near_cap, cap);
    p.near_counts = (int64_t *)counts;
    p.near_offsets = (const int64_t *)offsets;
    p.near_hist = (uint32_t *)hist;
    p.near_cursor = p.near_hist + ((size_t)p.n_keys + 2);
    p.near_state = (NearState *)state;
    p.near_cap = cap;
    p.near_keys = (uint32_t *)keys;
    p.near_recs = (uint64_t *)recs;
    p.near_vals = (double *)vals;
    p.near_width = width;
    L.scan_temp = temp;
    L.scan_bytes = tb;
    return CCOIL_OK;
}

// Checks the near pass's outcome after the call's work is done.
ccoil_status near_outcome(ccoil_ctx ctx, const NearState &st)
{
    if (st.mismatch)
        return fail(ctx, CCOIL_ERR_DEVICE, "FP32C near pass: the fill classified a (coil, point) pair differently "
                                           "from the field kernel; the near contributions are incomplete");
    return CCOIL_OK;
}

void enqueue_near_tail(hipStream_t s, const FieldLaunch &L, NearState *host_state, hipEvent_t evr0, hipEvent_t evr1)
{
    launch_near_key_scan(s, L.p);
    launch_near_fill(s, L.p);
    if (evr0) (void)hipEventRecord(evr0, s);
    launch_near_eval(s, L.p);
    if (evr1) (void)hipEventRecord(evr1, s);
    launch_near_add(s, L.p);
    if (host_state) (void)hipMemcpyAsync(host_state, L.p.near_state, sizeof(NearState), hipMemcpyDeviceToHost, s);
}

void enqueue_fields(hipStream_t s, const FieldLaunch &L, hipEvent_t evm0, hipEvent_t evm1, NearState *host_state,
                    hipEvent_t evr0, hipEvent_t evr1)
{
    const FieldParams &p = L.p;
    if (p.mp) launch_multipole_moments(s, p.coils, p.n_coils, p.tab, const_cast<double *>(p.mp));
    if (p.order)
        (void)launch_visit_order(s, p.pts, p.n_pts, L.order_bits, L.order_temp, L.order_scan_bytes,
                                 const_cast<uint32_t *>(p.order));
    if (L.near) {
        launch_fill_zero(s, p.near_counts, p.n_pts);
        launch_fill_zero(s, p.near_hist, ((int64_t)p.n_keys + 2) / 2);
        launch_fill_zero(s, p.near_state, sizeof(NearState) / 8);
    }
    if (evm0) (void)hipEventRecord(evm0, s);
    launch_field(s, p, L.precision, L.split);
    if (evm1) (void)hipEventRecord(evm1, s);
    launch_field_reduce(s, p);
    if (L.near) {
        size_t tb = L.scan_bytes;
        (void)near_scan(s, p.near_counts, const_cast<int64_t *>(p.near_offsets), p.n_pts, L.scan_temp, tb);
        if (!series_counts_keys(p)) launch_near_count(s, p);   // else the series kernel counted the keys
        enqueue_near_tail(s, L, host_state, evr0, evr1);
    }
}

ccoil_status fields_prepare(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                            uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B, double *E,
                            double *G, Arena *arena, const FieldShape *shape, int64_t near_cap, FieldPrep &out)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_fields: ctx is NULL");
    if (n_points < 0) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_points is negative");
    if (n_points > ((int64_t)1 << 31) - 1) return fail(ctx, CCOIL_ERR_ARGUMENT, "n_points exceeds 2^31 - 1");
    if (precision != CCOIL_FP64 && precision != CCOIL_FP32C && precision != CCOIL_FP32 &&
        precision != CCOIL_FP64_REFERENCE && precision != CCOIL_FP64_IEEE)
        return fail(ctx, CCOIL_ERR_ARGUMENT, "unknown precision");
    const uint32_t known = CCOIL_A | CCOIL_B | CCOIL_E | CCOIL_G | CCOIL_G_JACOBIAN;
    if (quantities == 0 || (quantities & ~known)) return fail(ctx, CCOIL_ERR_ARGUMENT, "bad quantity mask");
    if ((quantities & CCOIL_G) && (quantities & CCOIL_G_JACOBIAN))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "CCOIL_G and CCOIL_G_JACOBIAN share the G output; request one");
    if (((quantities & CCOIL_A) && !A) || ((quantities & CCOIL_B) && !B) || ((quantities & CCOIL_E) && !E) ||
        ((quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) && !G))
        return fail(ctx, CCOIL_ERR_ARGUMENT, "an output for a requested quantity is NULL");
    if (n_points > 0 && !points) return fail(ctx, CCOIL_ERR_ARGUMENT, "points is NULL");
    std::vector<CoilDev> &hc = out.hc;
    std::string msg = pack_coils(coils, false, (quantities & CCOIL_E) != 0, hc);
    if (!msg.empty()) return fail(ctx, CCOIL_ERR_ARGUMENT, msg);
    if (hc.size() > (size_t)INT32_MAX) return fail(ctx, CCOIL_ERR_ARGUMENT, "too many coils");
    CC_HIP(ctx, hipSetDevice(ctx->device));
    const bool dev = flags & CCOIL_DEVICE_POINTERS;
    if (arena && !dev) return fail(ctx, CCOIL_ERR_ARGUMENT, "a call without allocations takes device pointers only");
    const int64_t n = n_points;
    out.dev = dev;
    out.n = n;
    out.q = quantities;
    FieldLaunch &L = out.L;
    FieldParams &p = L.p;
    std::memset(&p, 0, sizeof p);
    p.n_pts = n;
    p.q = quantities;
    L.precision = precision;
    bool rule_records = false;
    L.near = field_record_pass(precision, flags, n_points, arena && arena->fixed, &rule_records);
    p.near_fp64 = precision == CCOIL_FP32C && !(flags & CCOIL_NO_NEAR_FP64);
    p.rule_records = rule_records;
    p.float_records = rule_records && precision == CCOIL_FP32C;
    const bool far = (precision == CCOIL_FP32C && !(flags & CCOIL_DIRECT)) ||
                     (precision == CCOIL_FP64 && (flags & CCOIL_FAR_FIELD));
    if (dev) {
        p.pts = points;
        p.A = (quantities & CCOIL_A) ? A : nullptr;
        p.B = (quantities & CCOIL_B) ? B : nullptr;
        p.E = (quantities & CCOIL_E) ? E : nullptr;
        p.G = (quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) ? G : nullptr;
    }
    out.width = ((quantities & CCOIL_A) ? 3 : 0) + ((quantities & CCOIL_E) ? 3 : 0) +
                ((quantities & CCOIL_B) ? 3 : 0) + ((quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) ? 9 : 0);
    out.nodes = 0;
    for (auto &c : hc) out.nodes += (int64_t)c.n_ang * c.n_thk;
    if (n == 0 || hc.empty()) return CCOIL_OK;   // nothing to evaluate, or every field zero
    const FieldShape sh = shape ? *shape : field_shape_rule(ctx, n, (int64_t)hc.size(), quantities, precision,
                                                            rule_records);
    if (sh.split != 1 && sh.split != 4 && sh.split != 16) return fail(ctx, CCOIL_ERR_ARGUMENT, "split must be 1, 4 or 16");
    L.split = rule_records ? 1 : sh.split;   // the series-only kernel runs one lane per point
    if (far)
        for (auto &c : hc) {
            // the cost floor holds for one lane per point; a launch that splits the rows runs the series on one
            // lane of the point's group, so it keeps the series from 2 r_max on
            const double y = far_distance(c, precision == CCOIL_FP64, L.split == 1);
            c.far_d2 = y > 0.0 ? (y * c.rmax) * (y * c.rmax) : 0.0;
            c.mp_terms = far_terms(y, precision == CCOIL_FP64);
        }

    // The FP32C near pass sorts its (coil, point) records by near_key: coils with the most nodes first, so its
    // longest waves start first. near_coil (near_key -> coil) follows the coils in the same buffer.
    std::vector<uint32_t> near_coil(hc.size());
    for (size_t i = 0; i < hc.size(); ++i) near_coil[i] = (uint32_t)i;
    auto cost = [&](uint32_t i) { return (int64_t)hc[i].n_ang * hc[i].n_thk; };   // field nodes of the coil
    std::stable_sort(near_coil.begin(), near_coil.end(), [&](uint32_t x, uint32_t y) { return cost(x) > cost(y); });
    for (size_t k = 0; k < hc.size(); ++k) hc[near_coil[k]].near_key = (int32_t)k;
    const size_t coil_bytes = sizeof(CoilDev) * hc.size();
    void *coil_mem = nullptr;
    ccoil_status st = call_buffer(ctx, ctx->coils, arena, coil_bytes + sizeof(uint32_t) * hc.size(), &coil_mem);
    if (st != CCOIL_OK) return st;
    CC_HIP(ctx, hipMemcpyAsync(coil_mem, hc.data(), coil_bytes, hipMemcpyHostToDevice, ctx->stream));
    CC_HIP(ctx, hipMemcpyAsync((char *)coil_mem + coil_bytes, near_coil.data(), sizeof(uint32_t) * hc.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    p.coils = (const CoilDev *)coil_mem;
    p.n_coils = (int32_t)hc.size();
    p.n_keys = p.n_coils * (p.float_records ? 2 : 1);
    p.near_coil = (const uint32_t *)((const char *)coil_mem + coil_bytes);
    p.tab = ctx_tables(ctx);
    if (!dev) {
        const size_t s3 = sizeof(double) * 3 * (size_t)n, s9 = sizeof(double) * 9 * (size_t)n;
        CC_HIP(ctx, ensure(ctx->pts, s3));
        CC_HIP(ctx, hipMemcpyAsync(ctx->pts.p, points, s3, hipMemcpyHostToDevice, ctx->stream));
        p.pts = (const double *)ctx->pts.p;
        if (quantities & CCOIL_A) { CC_HIP(ctx, ensure(ctx->outA, s3)); p.A = (double *)ctx->outA.p; }
        if (quantities & CCOIL_B) { CC_HIP(ctx, ensure(ctx->outB, s3)); p.B = (double *)ctx->outB.p; }
        if (quantities & CCOIL_E) { CC_HIP(ctx, ensure(ctx->outE, s3)); p.E = (double *)ctx->outE.p; }
        if (quantities & (CCOIL_G | CCOIL_G_JACOBIAN)) { CC_HIP(ctx, ensure(ctx->outG, s9)); p.G = (double *)ctx->outG.p; }
    }
    const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(sh.chunks, (int64_t)hc.size()));
    p.coils_per_chunk = (int32_t)(((int64_t)hc.size() + chunks - 1) / chunks);
    p.n_chunks = (int32_t)(((int64_t)hc.size() + p.coils_per_chunk - 1) / p.coils_per_chunk);
    if (p.n_chunks > 1) {
        p.partial_width = out.width;
        void *partial = nullptr;
        st = call_buffer(ctx, ctx->fpartial, arena, sizeof(double) * p.partial_width * (size_t)n * (size_t)p.n_chunks,
                         &partial);
        if (st != CCOIL_OK) return st;
        p.partial = (double *)partial;
    }
    if (far) {
        // the series' moments, and the visit order when there are enough points for the spatial sort to pay
        L.far_bytes = sizeof(double) * ((size_t)kMpHeader + (size_t)kMpCoil * hc.size()) + 256;
        void *mp = nullptr;
        st = call_buffer(ctx, ctx->fmp, arena, sizeof(double) * ((size_t)kMpHeader + (size_t)kMpCoil * hc.size()), &mp);
        if (st != CCOIL_OK) return st;
        p.mp = (const double *)mp;
        if (n >= kOrderMinPoints) {
            // about 16-32 points per cell: 2^bits cells per axis, bits = log8(n / 16), between 3 and 6
            int bits = 3;
            while (bits < 6 && ((int64_t)1 << (3 * (bits + 1))) * 16 <= n) ++bits;
            const int64_t cells = (int64_t)1 << (3 * bits);
            size_t sb = 0;
            CC_HIP(ctx, near_scan(ctx->stream, nullptr, nullptr, cells, nullptr, sb));
            const size_t ob = sizeof(uint32_t) * (size_t)n, tb = visit_order_bytes(n, bits, sb);
            void *order = nullptr, *temp = nullptr;
            if ((st = call_buffer(ctx, ctx->forder, arena, ob, &order)) != CCOIL_OK) return st;
            if ((st = call_buffer(ctx, ctx->ftemp, arena, tb, &temp)) != CCOIL_OK) return st;
            L.far_bytes += ob + tb + 2 * 256;
            p.order = (const uint32_t *)order;
            L.order_temp = temp;
            L.order_scan_bytes = sb;
            L.order_bits = bits;
        }
    }
    if (L.near) {
        // Buffers for the records: every (coil, point) pair when that is small, else a capacity kept from earlier
        // calls (at least two per point); a call that finds more grows them and runs the near pass again.
        const int64_t all_pairs = n * (int64_t)hc.size();
        int64_t cap = near_cap > 0 ? near_cap
                                   : std::max<int64_t>(ctx->near_cap, std::min<int64_t>(all_pairs, 2 * n + 65536));
        cap = std::min<int64_t>(std::min<int64_t>(cap, all_pairs), INT32_MAX);
        st = near_buffers(ctx, L, cap, out.width, arena);
        if (st != CCOIL_OK) return st;
    }
    return CCOIL_OK;
}

// Zero-fills the requested outputs of a call with no coils (the reference returns a zero-filled array).
void enqueue_zero_outputs(hipStream_t s, const FieldPrep &fp)
{
    const FieldParams &p = fp.L.p;
    double *outs[4] = {p.A, p.B, p.E, p.G};
    const int64_t words[4] = {3 * fp.n, 3 * fp.n, 3 * fp.n, 9 * fp.n};
    for (int k = 0; k < 4; ++k)
        if (outs[k]) launch_fill_zero(s, outs[k], words[k]);
}

ccoil_status fields_call(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                         uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B, double *E,
                         double *G, ccoil_report *report, Arena *arena, const FieldShape *shape)
{
    if (report) std::memset(report, 0, sizeof *report);
    FieldPrep fp;
    ccoil_status st = fields_prepare(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G, arena,
                                     shape, 0, fp);
    if (st != CCOIL_OK) return st;
    const int64_t n = fp.n;
    const size_t s3 = sizeof(double) * 3 * (size_t)n, s9 = sizeof(double) * 9 * (size_t)n;
    if (n == 0) return CCOIL_OK;
    if (fp.hc.empty()) {
        if (fp.dev) {
            enqueue_zero_outputs(ctx->stream, fp);
            CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        } else {
            double *outs[4] = {A, B, E, G};
            const uint32_t bits[4] = {CCOIL_A, CCOIL_B, CCOIL_E, CCOIL_G | CCOIL_G_JACOBIAN};
            const size_t sz[4] = {s3, s3, s3, s9};
            for (int k = 0; k < 4; ++k)
                if (quantities & bits[k]) std::memset(outs[k], 0, sz[k]);
        }
        return CCOIL_OK;
    }
    FieldLaunch &L = fp.L;
    const FieldParams &p = L.p;
    CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    enqueue_fields(ctx->stream, L, ctx->evm0, ctx->evm1, L.near ? ctx->near_host : nullptr, ctx->evr0, ctx->evr1);
    CC_HIP(ctx, hipGetLastError());
    CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    auto copy_out = [&]() -> ccoil_status {
        if (fp.dev) return CCOIL_OK;
        if (p.A) CC_HIP(ctx, hipMemcpyAsync(A, p.A, s3, hipMemcpyDeviceToHost, ctx->stream));
        if (p.B) CC_HIP(ctx, hipMemcpyAsync(B, p.B, s3, hipMemcpyDeviceToHost, ctx->stream));
        if (p.E) CC_HIP(ctx, hipMemcpyAsync(E, p.E, s3, hipMemcpyDeviceToHost, ctx->stream));
        if (p.G) CC_HIP(ctx, hipMemcpyAsync(G, p.G, s9, hipMemcpyDeviceToHost, ctx->stream));
        return CCOIL_OK;
    };
    st = copy_out();
    if (st != CCOIL_OK) return st;
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t near = 0, total = 0;
    if (L.near) {
        st = near_outcome(ctx, *ctx->near_host);
        if (st != CCOIL_OK) return st;
        total = (int64_t)ctx->near_host->total;
        near = (int64_t)ctx->near_host->fp64;
        if (ctx->near_host->overflow && arena)
            return fail(ctx, CCOIL_ERR_ARGUMENT, "FP32C: " + std::to_string(total) + " near-winding (coil, point) "
                        "pairs exceed the " + std::to_string(L.p.near_cap) + " records the fixed scratch holds for "
                        "this call; the near contributions are missing; split the points");
        if (ctx->near_host->overflow) {
            // more records than the buffers hold: the pass did nothing; grow and run it again
            st = near_buffers(ctx, L, std::min<int64_t>(total + total / 8 + 4096, INT32_MAX), fp.width, nullptr);
            if (st != CCOIL_OK) return st;
            enqueue_near_tail(ctx->stream, L, ctx->near_host, ctx->evr0, ctx->evr1);
            CC_HIP(ctx, hipGetLastError());
            CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
            st = copy_out();
            if (st != CCOIL_OK) return st;
            CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            st = near_outcome(ctx, *ctx->near_host);
            if (st != CCOIL_OK) return st;
            if (ctx->near_host->overflow)
                return fail(ctx, CCOIL_ERR_DEVICE, "record pass: capacity still exceeded after growing");
            near = (int64_t)ctx->near_host->fp64;
        }
        if (precision != CCOIL_FP32C) near = 0;   // FP64's records are its rule, not a re-evaluation
    }
    if (report) {
        float ms = 0.f, mms = 0.f;
        CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        CC_HIP(ctx, hipEventElapsedTime(&mms, ctx->evm0, ctx->evm1));
        report->node_evaluations = fp.nodes * n;
        report->near_pairs = near;
        report->kernel_ms = ms;
        report->main_kernel_ms = mms;
        if (L.near) {
            float rms = 0.f;
            CC_HIP(ctx, hipEventElapsedTime(&rms, ctx->evr0, ctx->evr1));
            report->record_ms = rms;
        }
        report->block_size = kFieldBlock;
        report->split = L.split;
    }
    return CCOIL_OK;
}

}  // namespace ccoil

extern "C" ccoil_status ccoil_far_field_radius(const ccoil_coils *coils, int32_t precision, double *rmax,
                                               double *distance)
{
    std::vector<CoilDev> hc;
    const std::string msg = pack_coils(coils, false, false, hc);
    if (!msg.empty()) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_far_field_radius: " + msg);
    if (!hc.empty() && (!rmax || !distance))
        return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_far_field_radius: an output is NULL");
    const bool far = precision == CCOIL_FP32C || precision == CCOIL_FP64;
    host_parallel_for((int64_t)hc.size(), [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            const CoilDev &c = hc[(size_t)i];
            rmax[i] = c.rmax;
            distance[i] = far ? far_distance(c, precision == CCOIL_FP64, true) * c.rmax : 0.0;
        }
    });
    return CCOIL_OK;
}

extern "C" ccoil_status ccoil_fields(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                                     uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B,
                                     double *E, double *G, ccoil_report *report)
{
    const FieldShape *shape = nullptr;
    FieldShape tuned;
    if (ctx && (flags & CCOIL_AUTOTUNE)) {
        const ccoil_status st = tune_fields(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G,
                                            tuned);
        if (st != CCOIL_OK) return st;
        shape = &tuned;
    }
    // A series-mode call too large for one record pass runs as pieces of points, each with its own pass (records
    // grow with points x coils). Each piece is a call of its own: its launch shape follows its size, so a point in
    // a short last piece can differ in the last bits from the same point in a longer one
    const bool far = (precision == CCOIL_FP32C && !(flags & CCOIL_DIRECT)) ||
                     (precision == CCOIL_FP64 && (flags & CCOIL_FAR_FIELD));
    const int64_t n_coils = coils ? coils->n : 0;
    const int64_t piece = n_coils > 0 ? std::max<int64_t>(kOrderMinPoints, std::min<int64_t>(kRecordMaxPoints,
                                            kRecordMaxPoints * 128 / n_coils)) : kRecordMaxPoints;
    if (!ctx || !far || shape || n_points <= piece || !points)
        return fields_call(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G, report, nullptr,
                           shape);
    if (report) std::memset(report, 0, sizeof *report);
    for (int64_t s0 = 0; s0 < n_points; s0 += piece) {
        const int64_t m = std::min<int64_t>(piece, n_points - s0);
        ccoil_report r;
        const ccoil_status st = fields_call(ctx, coils, points + 3 * s0, m, quantities, precision, flags,
                                            A ? A + 3 * s0 : nullptr, B ? B + 3 * s0 : nullptr,
                                            E ? E + 3 * s0 : nullptr, G ? G + 9 * s0 : nullptr, &r, nullptr, nullptr);
        if (st != CCOIL_OK) return st;
        if (report) {
            report->node_evaluations += r.node_evaluations;
            report->near_pairs += r.near_pairs;
            report->kernel_ms += r.kernel_ms;
            report->main_kernel_ms += r.main_kernel_ms;
            report->record_ms += r.record_ms;
            report->block_size = r.block_size;
            report->split = r.split;
        }
    }
    return CCOIL_OK;
}
