#include "GpuBridge.h"

#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

namespace ccoil {
namespace bridge {

namespace {

struct CtxHolder {
    ccoil_ctx ctx = nullptr;
    ~CtxHolder()
    {
        if (ctx) ccoil_destroy(ctx);
    }
};

thread_local CtxHolder t_ctx;
thread_local int t_device = 0;
thread_local int t_precision = CCOIL_FP32C;
thread_local int t_far_field = -1;

}  // namespace

ccoil_ctx context()
{
    if (!t_ctx.ctx) {
        int dev = 0;
        if (const char *e = std::getenv("CCOIL_DEVICE")) dev = std::atoi(e);
        ccoil_ctx c = nullptr;
        ccoil_status st = ccoil_create(dev, &c);
        if (st != CCOIL_OK) check(st, nullptr, "ccoil_create");
        t_ctx.ctx = c;
        t_device = dev;
    }
    return t_ctx.ctx;
}

void check(ccoil_status st, ccoil_ctx ctx, const char *what)
{
    if (st == CCOIL_OK) return;
    const std::string msg = std::string(what) + ": " + ccoil_last_error(ctx);
    switch (st) {
        case CCOIL_ERR_ARGUMENT: throw std::invalid_argument(msg);
        case CCOIL_ERR_NO_DEVICE: throw std::logic_error("GPU functions are unavailable: " + msg);
        case CCOIL_ERR_OUT_OF_MEMORY: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

int device()
{
    (void)context();
    return t_device;
}

void *stream()
{
    ccoil_ctx ctx = context();
    void *s = nullptr;
    check(ccoil_get_stream(ctx, &s), ctx, "ccoil_get_stream");
    return s;
}

int precision() { return t_precision; }

void setPrecision(int p)
{
    if (p != CCOIL_FP64 && p != CCOIL_FP32C && p != CCOIL_FP32 && p != CCOIL_FP64_REFERENCE &&
        p != CCOIL_FP64_IEEE)
        throw std::invalid_argument("GPU precision must be CCOIL_FP64, CCOIL_FP32C, CCOIL_FP32, "
                                    "CCOIL_FP64_REFERENCE or CCOIL_FP64_IEEE");
    t_precision = p;
}

int farField() { return t_far_field; }

void setFarField(int mode)
{
    if (mode < -1 || mode > 1) throw std::invalid_argument("far-field mode must be -1 (default), 0 (off) or 1 (on)");
    t_far_field = mode;
}

uint32_t fieldFlags()
{
    if (t_far_field < 0) return 0u;
    if (t_precision == CCOIL_FP32C) return t_far_field ? 0u : (uint32_t)CCOIL_DIRECT;
    if (t_precision == CCOIL_FP64) return t_far_field ? (uint32_t)CCOIL_FAR_FIELD : 0u;
    return 0u;
}

void CoilSoA::add(const Coil &c, const PrecisionArguments &args)
{
    auto ang = c.getRotationAngles();
    add(c, args, c.getPositionVector(), ang.first, ang.second);
}

void CoilSoA::add(const Coil &c, const PrecisionArguments &args, const vec3::Vector3 &position, double yAngle,
                  double zAngle)
{
    use_fast.push_back(c.isUsingFastMethod() ? 1 : 0);
    n_ang.push_back(args.angularIncrements);
    n_thk.push_back(args.thicknessIncrements);
    n_len.push_back(args.lengthIncrements);
    blocks.push_back(args.angularBlocks);
    blocks.push_back(args.thicknessBlocks);
    blocks.push_back(args.lengthBlocks);
    R.push_back(c.getInnerRadius());
    a.push_back(c.getThickness());
    b.push_back(c.getLength());
    J.push_back(c.getCurrentDensity());
    current.push_back(c.getCurrent());
    turns.push_back((double)c.getNumOfTurns());
    freq.push_back(c.getSineFrequency());
    pos.push_back(position.x);
    pos.push_back(position.y);
    pos.push_back(position.z);
    alpha.push_back(yAngle);
    beta.push_back(zAngle);
}

ccoil_coils CoilSoA::view() const
{
    ccoil_coils v;
    v.n = (int64_t)R.size();
    v.use_fast = use_fast.data();
    v.n_ang = n_ang.data();
    v.n_thk = n_thk.data();
    v.n_len = n_len.data();
    v.blocks = blocks.data();
    v.R = R.data();
    v.a = a.data();
    v.b = b.data();
    v.J = J.data();
    v.current = current.data();
    v.turns = turns.data();
    v.frequency = freq.data();
    v.pos = pos.data();
    v.alpha = alpha.data();
    v.beta = beta.data();
    return v;
}

std::vector<std::pair<vec3::Vector3, vec3::Vector3>> dipoleForceTorque(const CoilSoA &coils,
                                                                       const vec3::Vector3Array &points,
                                                                       const vec3::Vector3Array &moments)
{
    if (points.size() != moments.size()) throw std::logic_error("Array sizes do not match!");
    const size_t n = points.size();
    std::vector<double> B(3 * n), J(9 * n);
    const ccoil_coils v = coils.view();
    ccoil_ctx ctx = context();
    check(ccoil_fields(ctx, &v, points.data(), (int64_t)n, CCOIL_B | CCOIL_G_JACOBIAN, precision(), fieldFlags(),
                       nullptr,
                       B.data(), nullptr, J.data(), nullptr),
          ctx, "ccoil_fields");
    std::vector<std::pair<vec3::Vector3, vec3::Vector3>> out(n);
    for (size_t k = 0; k < n; ++k) {
        const vec3::Vector3 m = moments[k];
        const double *j = &J[9 * k], *b = &B[3 * k];
        const vec3::Vector3 F(j[0] * m.x + j[3] * m.y + j[6] * m.z, j[1] * m.x + j[4] * m.y + j[7] * m.z,
                              j[2] * m.x + j[5] * m.y + j[8] * m.z);
        const vec3::Vector3 T(m.y * b[2] - m.z * b[1], m.z * b[0] - m.x * b[2], m.x * b[1] - m.y * b[0]);
        out[k] = {F, T};
    }
    return out;
}

void fieldsInto(const CoilSoA &coils, const double *points, size_t n, double *A, double *B, double *E, double *G,
                bool devicePointers)
{
    const uint32_t q = (A ? CCOIL_A : 0u) | (B ? CCOIL_B : 0u) | (E ? CCOIL_E : 0u) | (G ? CCOIL_G : 0u);
    if (!q) throw std::invalid_argument("no output requested: pass at least one of A, B, E, G");
    const ccoil_coils v = coils.view();
    ccoil_ctx ctx = context();
    check(ccoil_fields(ctx, &v, points, (int64_t)n, q, precision(),
                       fieldFlags() | (devicePointers ? (uint32_t)CCOIL_DEVICE_POINTERS : 0u), A, B,
                       E, G, nullptr),
          ctx, "ccoil_fields");
}

namespace {

bool fastPair(const Coil &c)
{
    return c.getCoilType() == CoilType::THIN || c.getCoilType() == CoilType::RECTANGULAR;
}

// Gauss-Legendre rule of order n as the device generates it (cached per thread).
const std::pair<std::vector<double>, std::vector<double>> &rule(int n)
{
    thread_local std::map<int, std::pair<std::vector<double>, std::vector<double>>> cache;
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    std::vector<double> x((size_t)n), w((size_t)n);
    ccoil_ctx ctx = context();
    check(ccoil_gauss_legendre(ctx, n, x.data(), w.data()), ctx, "ccoil_gauss_legendre");
    return cache.emplace(n, std::make_pair(std::move(x), std::move(w))).first->second;
}

// The slow coaxial rule: quantity q (CCOIL_A or CCOIL_B) of the primary at the secondary's r-z grid, weighted by
// 0.25 r w_z w_r, summed in grid order and divided by the number of blocks.
double zAxisSlow(const Coil &prim, const Coil &sec, const CoilPairArguments &args, uint32_t q)
{
    const PrecisionArguments &sp = args.secondaryPrecision;
    const int lb = sp.lengthBlocks, li = sp.lengthIncrements, tb = sp.thicknessBlocks, ti = sp.thicknessIncrements;
    const double lbs = sec.getLength() / lb, tbs = sec.getThickness() / tb;
    const auto &rz = rule(li);
    const auto &rr = rule(ti);
    const double zDisp = sec.getPositionVector().z;
    vec3::Vector3Array pts;
    std::vector<double> w;
    pts.reserve((size_t)lb * li * tb * ti);
    w.reserve((size_t)lb * li * tb * ti);
    for (int zb = 0; zb < lb; ++zb)
        for (int rb = 0; rb < tb; ++rb) {
            const double zbp = -1 * (sec.getLength() * 0.5) + lbs * (zb + 0.5);
            const double rbp = sec.getInnerRadius() + tbs * (rb + 0.5);
            for (int zi = 0; zi < li; ++zi)
                for (int ri = 0; ri < ti; ++ri) {
                    const double z = zDisp + zbp + (lbs * 0.5) * rz.first[(size_t)zi];
                    const double r = rbp + (tbs * 0.5) * rr.first[(size_t)ri];
                    pts.append(r, 0.0, z);
                    w.push_back(0.25 * r * rz.second[(size_t)zi] * rr.second[(size_t)ri]);
                }
        }
    CoilSoA p;
    p.add(prim, args.primaryPrecision);
    const ccoil_coils v = p.view();
    std::vector<double> out(3 * pts.size());
    ccoil_ctx ctx = context();
    check(ccoil_fields(ctx, &v, pts.data(), (int64_t)pts.size(), q, precision(), fieldFlags(),
                       q == CCOIL_A ? out.data() : nullptr,
                       q == CCOIL_B ? out.data() : nullptr, nullptr, nullptr, nullptr),
          ctx, "ccoil_fields");
    double sum = 0.0;
    for (size_t i = 0; i < pts.size(); ++i) {
        const double *o = &out[3 * i];
        sum += (q == CCOIL_A ? std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) : o[0]) * w[i];
    }
    return sum / (lb * tb);
}

}  // namespace

double zAxisMutualInductance(const Coil &prim, const Coil &sec, const CoilPairArguments &args)
{
    if (fastPair(prim) && fastPair(sec)) {
        CoilSoA p, s;
        p.add(prim, args.primaryPrecision);
        s.add(sec, args.secondaryPrecision);
        const ccoil_coils pv = p.view(), sv = s.view();
        double M = 0.0;
        ccoil_ctx ctx = context();
        check(ccoil_pairs_zaxis(ctx, &pv, &sv, nullptr, nullptr, 1, CCOIL_PAIR_M, 0, &M, nullptr, nullptr), ctx,
              "ccoil_pairs_zaxis");
        return M;
    }
    return zAxisSlow(prim, sec, args, CCOIL_A) * 2 * M_PI * sec.getNumOfTurns() / prim.getCurrent();
}

double zAxisForce(const Coil &prim, const Coil &sec, const CoilPairArguments &args)
{
    if (fastPair(prim) && fastPair(sec)) {
        CoilSoA p, s;
        p.add(prim, args.primaryPrecision);
        s.add(sec, args.secondaryPrecision);
        const ccoil_coils pv = p.view(), sv = s.view();
        double F = 0.0;
        ccoil_ctx ctx = context();
        check(ccoil_pairs_zaxis(ctx, &pv, &sv, nullptr, nullptr, 1, CCOIL_PAIR_FT, 0, nullptr, &F, nullptr), ctx,
              "ccoil_pairs_zaxis");
        return F;
    }
    return (-1.0) * zAxisSlow(prim, sec, args, CCOIL_B) * 2 * M_PI * sec.getNumOfTurns() * sec.getCurrent();
}

double selfInductance(const Coil &coil, const CoilPairArguments &args)
{
    CoilSoA c;
    c.add(coil, args.primaryPrecision);
    const ccoil_coils v = c.view();
    const int32_t radial[2] = {args.secondaryPrecision.thicknessBlocks, args.secondaryPrecision.thicknessIncrements};
    double L = 0.0;
    ccoil_ctx ctx = context();
    check(ccoil_self_inductance(ctx, &v, radial, 0, &L, nullptr), ctx, "ccoil_self_inductance");
    return L;
}

}  // namespace bridge
}  // namespace ccoil
