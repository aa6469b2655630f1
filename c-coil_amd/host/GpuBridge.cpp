#include "GpuBridge.h"

#include <cstdlib>
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
thread_local int t_precision = CCOIL_FP32C;

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

int precision() { return t_precision; }

void setPrecision(int p)
{
    if (p != CCOIL_FP64 && p != CCOIL_FP32C && p != CCOIL_FP32 && p != CCOIL_FP64_REFERENCE &&
        p != CCOIL_FP64_IEEE)
        throw std::invalid_argument("GPU precision must be CCOIL_FP64, CCOIL_FP32C, CCOIL_FP32, "
                                    "CCOIL_FP64_REFERENCE or CCOIL_FP64_IEEE");
    t_precision = p;
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
    check(ccoil_fields(ctx, &v, points.data(), (int64_t)n, CCOIL_B | CCOIL_G_JACOBIAN, precision(), 0, nullptr,
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

}  // namespace bridge
}  // namespace ccoil
