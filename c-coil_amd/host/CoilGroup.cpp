// CoilGroup: container semantics of src/CoilGroup/CoilGroup.cxx:20-114; GPU paths of
// src/CoilGroup/Fields/CalculateFieldsGPU.cxx:11-194 (all members at the group's PrecisionFactor,
// GPUArgumentGeneration.cxx:17-95) and src/CoilGroup/{MInductance,ForceAndTorque}/Calculate*GPU.cxx
// (PrecisionFactor remapped to (pf - 1)/2 + 1, secondary planned with getSecondaryCoilPrecisionArgumentsGPU).
#include "CoilGroup.h"
#include "GpuBridge.h"

#include <sstream>
#include <stdexcept>

using ccoil::bridge::CoilSoA;

namespace {

[[noreturn]] void cpuOnly(const char *what)
{
    throw std::logic_error(std::string(what) +
                           " is computed on the CPU by C-Coil; this library provides the GPU path "
                           "(use the computeAll* / arrangement methods with ComputeMethod::GPU)");
}

}  // namespace

CoilGroup::CoilGroup(std::vector<std::shared_ptr<Coil>> coils, PrecisionFactor pf, int threads)
    : memberCoils(std::move(coils)), defaultPrecisionFactor(pf), threadCount(threads) {}

PrecisionFactor CoilGroup::getDefaultPrecisionFactor() const { return defaultPrecisionFactor; }
int CoilGroup::getThreadCount() const { return threadCount; }
const std::vector<std::shared_ptr<Coil>> &CoilGroup::getMemberCoils() const { return memberCoils; }

void CoilGroup::setDefaultPrecisionFactor(PrecisionFactor pf)
{
    defaultPrecisionFactor = pf;
    for (auto &c : memberCoils) c->setDefaultPrecision(pf);
}

void CoilGroup::setThreadCount(int t)
{
    threadCount = t;
    for (auto &c : memberCoils) c->setThreadCount(t);
}

void CoilGroup::addCoil(double R, double a, double b, int N, double I, PrecisionFactor pf, int threads,
                        vec3::Vector3 pos, double yA, double zA)
{
    memberCoils.push_back(std::make_shared<Coil>(R, a, b, N, I, pf, threads, pos, yA, zA));
}

void CoilGroup::removeCoil(size_t index)
{
    if (index >= memberCoils.size()) throw std::out_of_range("Coil index out of range!");
    memberCoils.erase(memberCoils.begin() + (long)index);
}

Coil &CoilGroup::operator[](size_t index) const
{
    if (index >= memberCoils.size()) throw std::out_of_range("Coil index out of range!");
    return *memberCoils[index];
}

bool CoilGroup::isPointInside(vec3::Vector3 p)
{
    for (auto &c : memberCoils)
        if (c->isPointInside(p)) return true;
    return false;
}

vec3::Vector3 CoilGroup::computeAPotentialVector(vec3::Vector3) const { cpuOnly("CoilGroup::computeAPotentialVector"); }
vec3::Vector3 CoilGroup::computeBFieldVector(vec3::Vector3) const { cpuOnly("CoilGroup::computeBFieldVector"); }
vec3::Vector3 CoilGroup::computeEFieldVector(vec3::Vector3) const { cpuOnly("CoilGroup::computeEFieldVector"); }
vec3::Matrix3 CoilGroup::computeBGradientMatrix(vec3::Vector3) const { cpuOnly("CoilGroup::computeBGradientMatrix"); }
std::pair<vec3::Vector3, vec3::Vector3> CoilGroup::computeForceOnDipoleMoment(vec3::Vector3, vec3::Vector3) const
{
    cpuOnly("CoilGroup::computeForceOnDipoleMoment");
}

namespace {

void groupFields(const std::vector<std::shared_ptr<Coil>> &coils, PrecisionFactor pf, const vec3::Vector3Array &pts,
                 uint32_t q, double *out)
{
    CoilSoA s;
    for (auto &c : coils) s.add(*c, PrecisionArguments::getCoilPrecisionArgumentsGPU(*c, pf));
    const ccoil_coils v = s.view();
    ccoil_ctx ctx = ccoil::bridge::context();
    double *A = (q & CCOIL_A) ? out : nullptr, *B = (q & CCOIL_B) ? out : nullptr;
    double *E = (q & CCOIL_E) ? out : nullptr, *G = (q & CCOIL_G) ? out : nullptr;
    ccoil::bridge::check(ccoil_fields(ctx, &v, pts.data(), (int64_t)pts.size(), q, ccoil::bridge::precision(), 0, A,
                                      B, E, G, nullptr),
                         ctx, "ccoil_fields");
}

}  // namespace

void CoilGroup::computeAllFieldsGPU(const double *points, size_t n, double *A, double *B, double *E, double *G,
                                    bool devicePointers) const
{
    CoilSoA s;
    for (auto &c : memberCoils) s.add(*c, PrecisionArguments::getCoilPrecisionArgumentsGPU(*c, defaultPrecisionFactor));
    ccoil::bridge::fieldsInto(s, points, n, A, B, E, G, devicePointers);
}

std::vector<std::pair<vec3::Vector3, vec3::Vector3>>
CoilGroup::computeAllForceOnDipoleMoments(const vec3::Vector3Array &points, const vec3::Vector3Array &moments,
                                          ComputeMethod m) const
{
    if (m != GPU) cpuOnly("CoilGroup::computeAllForceOnDipoleMoments with CPU_ST/CPU_MT");
    CoilSoA s;
    for (auto &c : memberCoils) s.add(*c, PrecisionArguments::getCoilPrecisionArgumentsGPU(*c, defaultPrecisionFactor));
    return ccoil::bridge::dipoleForceTorque(s, points, moments);
}

vec3::Vector3Array CoilGroup::computeAllAPotentialVectors(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    if (m != GPU) cpuOnly("CoilGroup::computeAllAPotentialVectors with CPU_ST/CPU_MT");
    vec3::Vector3Array out(pts.size());
    groupFields(memberCoils, defaultPrecisionFactor, pts, CCOIL_A, out.data());
    return out;
}

vec3::Vector3Array CoilGroup::computeAllBFieldVectors(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    if (m != GPU) cpuOnly("CoilGroup::computeAllBFieldVectors with CPU_ST/CPU_MT");
    vec3::Vector3Array out(pts.size());
    groupFields(memberCoils, defaultPrecisionFactor, pts, CCOIL_B, out.data());
    return out;
}

vec3::Vector3Array CoilGroup::computeAllEFieldVectors(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    if (m != GPU) cpuOnly("CoilGroup::computeAllEFieldVectors with CPU_ST/CPU_MT");
    vec3::Vector3Array out(pts.size());
    groupFields(memberCoils, defaultPrecisionFactor, pts, CCOIL_E, out.data());
    return out;
}

vec3::Matrix3Array CoilGroup::computeAllBGradientMatrices(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    if (m != GPU) cpuOnly("CoilGroup::computeAllBGradientMatrices with CPU_ST/CPU_MT");
    vec3::Matrix3Array out(pts.size());
    groupFields(memberCoils, defaultPrecisionFactor, pts, CCOIL_G, out.data());
    return out;
}

namespace {

// Every member except `secondary` (matched by id, CoilGroup/MInductance/ComputeMInductance.cxx:14-16) against
// the secondary, each pair planned and dispatched as Coil::computeMutualInductance / computeForceTorque do it:
// coaxial members take the z-axis rules (the THIN/RECTANGULAR ones batched in one ccoil_pairs_zaxis call), all
// others the general rule (one ccoil_pairs call). Member results are summed in member order.
void groupPairsGeneral(const std::vector<std::shared_ptr<Coil>> &coils, const Coil &sec, PrecisionFactor pf,
                       uint32_t what, double *M, double *F, double *T)
{
    CoilSoA p, s, zp, zs;
    std::vector<double> offsets;
    std::vector<int32_t> halves;
    enum Kind { kGeneral, kZFast, kZSlow };
    std::vector<std::pair<Kind, size_t>> members;   // (rule, index within that rule's batch)
    std::vector<double> slowM, slowF;
    auto fast = [](const Coil &c) {
        return c.getCoilType() == CoilType::THIN || c.getCoilType() == CoilType::RECTANGULAR;
    };
    for (auto &c : coils) {
        if (c->getId() == sec.getId()) continue;
        const bool z = Coil::isZAxisCase(*c, sec);
        const auto args = CoilPairArguments::getAppropriateCoilPairArguments(*c, sec, pf, GPU, z, false);
        if (z && fast(*c) && fast(sec)) {
            members.emplace_back(kZFast, zp.size());
            zp.add(*c, args.primaryPrecision);
            zs.add(sec, args.secondaryPrecision);
        } else if (z) {
            members.emplace_back(kZSlow, slowM.size());
            slowM.push_back((what & CCOIL_PAIR_M) ? ccoil::bridge::zAxisMutualInductance(*c, sec, args) : 0.0);
            slowF.push_back((what & CCOIL_PAIR_FT) ? ccoil::bridge::zAxisForce(*c, sec, args) : 0.0);
        } else {
            members.emplace_back(kGeneral, p.size());
            p.add(*c, args.primaryPrecision);
            s.add(sec, args.secondaryPrecision);
            auto ip = Coil::improvedPrecisionCase(*c, sec);
            const bool useHalf = (what & CCOIL_PAIR_M) && ip.first;
            offsets.push_back(useHalf ? ip.second : 0.0);
            halves.push_back(useHalf ? 1 : 0);
        }
    }
    if (what & CCOIL_PAIR_M) *M = 0.0;
    if (what & CCOIL_PAIR_FT)
        for (int k = 0; k < 3; ++k) F[k] = T[k] = 0.0;
    const int64_t n = (int64_t)p.size(), nz = (int64_t)zp.size();
    std::vector<double> m(n), f(3 * n), t(3 * n), zm(nz), zf(nz);
    ccoil_ctx ctx = ccoil::bridge::context();
    if (n > 0) {
        const ccoil_coils pv = p.view(), sv = s.view();
        ccoil::bridge::check(ccoil_pairs(ctx, &pv, &sv, nullptr, nullptr, n, offsets.data(), halves.data(), what,
                                         ccoil::bridge::precision(), 0, m.data(), f.data(), t.data(), nullptr,
                                         nullptr),
                             ctx, "ccoil_pairs");
    }
    if (nz > 0) {
        const ccoil_coils pv = zp.view(), sv = zs.view();
        ccoil::bridge::check(ccoil_pairs_zaxis(ctx, &pv, &sv, nullptr, nullptr, nz, what & (CCOIL_PAIR_M | CCOIL_PAIR_FT),
                                               0, zm.data(), zf.data(), nullptr),
                             ctx, "ccoil_pairs_zaxis");
    }
    for (const auto &mk : members) {
        const size_t i = mk.second;
        if (mk.first == kGeneral) {
            if (what & CCOIL_PAIR_M) *M += m[i];
            if (what & CCOIL_PAIR_FT)
                for (int k = 0; k < 3; ++k) {
                    F[k] += f[3 * i + k];
                    T[k] += t[3 * i + k];
                }
        } else {
            // a coaxial pair has only the axial force and no torque (ComputeForceMethods.cxx:8-21)
            if (what & CCOIL_PAIR_M) *M += mk.first == kZFast ? zm[i] : slowM[i];
            if (what & CCOIL_PAIR_FT) F[2] += mk.first == kZFast ? zf[i] : slowF[i];
        }
    }
}

}  // namespace

double CoilGroup::computeMutualInductance(const Coil &sec, PrecisionFactor pf, ComputeMethod m) const
{
    if (m != GPU) cpuOnly("CoilGroup::computeMutualInductance with CPU_ST/CPU_MT");
    double M = 0.0;
    groupPairsGeneral(memberCoils, sec, pf, CCOIL_PAIR_M, &M, nullptr, nullptr);
    return M;
}

std::pair<vec3::Vector3, vec3::Vector3> CoilGroup::computeForceTorque(const Coil &sec, PrecisionFactor pf,
                                                                      ComputeMethod m) const
{
    if (m != GPU) cpuOnly("CoilGroup::computeForceTorque with CPU_ST/CPU_MT");
    double F[3], T[3];
    groupPairsGeneral(memberCoils, sec, pf, CCOIL_PAIR_FT, nullptr, F, T);
    return {vec3::Vector3(F[0], F[1], F[2]), vec3::Vector3(T[0], T[1], T[2])};
}

namespace {

// CoilGroup pure-GPU arrangements: members x configurations pairs, summed per configuration
// (MInductanceGroup.cu:10-143, ForceGroup.cu:10-159; host CalculateMInductanceGPU.cxx:11-84).
void groupArrangements(const std::vector<std::shared_ptr<Coil>> &coils, const Coil &sec,
                       const vec3::Vector3Array &sp, const std::vector<double> &sY, const std::vector<double> &sZ,
                       PrecisionFactor pf, uint32_t what, std::vector<double> &M, std::vector<double> &F,
                       std::vector<double> &T)
{
    const size_t nc = sp.size();
    if (nc != sY.size() || nc != sZ.size()) throw std::logic_error("Array sizes do not match!");
    const PrecisionFactor gpf(((pf.relativePrecision - 1.0) / 2.0) + 1.0);
    const PrecisionArguments secArgs = PrecisionArguments::getSecondaryCoilPrecisionArgumentsGPU(sec, gpf);
    CoilSoA p, s;
    for (auto &c : coils)
        if (c->getId() != sec.getId()) p.add(*c, PrecisionArguments::getCoilPrecisionArgumentsGPU(*c, gpf));
    for (size_t i = 0; i < nc; ++i) s.add(sec, secArgs, sp[i], sY[i], sZ[i]);
    const int64_t np = (int64_t)p.size();
    M.assign(nc, 0.0);
    F.assign(3 * nc, 0.0);
    T.assign(3 * nc, 0.0);
    if (np == 0 || nc == 0) return;
    const int64_t n = np * (int64_t)nc;
    std::vector<int64_t> pi(n), si(n);
    for (int64_t k = 0; k < n; ++k) {
        pi[k] = k % np;
        si[k] = k / np;
    }
    std::vector<double> m(n), f(3 * n), t(3 * n);
    const ccoil_coils pv = p.view(), sv = s.view();
    ccoil_ctx ctx = ccoil::bridge::context();
    ccoil::bridge::check(ccoil_pairs(ctx, &pv, &sv, pi.data(), si.data(), n, nullptr, nullptr, what,
                                     ccoil::bridge::precision(), 0, m.data(), f.data(), t.data(), nullptr, nullptr),
                         ctx, "ccoil_pairs");
    for (int64_t k = 0; k < n; ++k) {
        const size_t c = (size_t)si[k];
        if (what & CCOIL_PAIR_M) M[c] += m[k];
        if (what & CCOIL_PAIR_FT)
            for (int q = 0; q < 3; ++q) {
                F[3 * c + q] += f[3 * k + q];
                T[3 * c + q] += t[3 * k + q];
            }
    }
}

}  // namespace

std::vector<double> CoilGroup::computeAllMutualInductanceArrangements(const Coil &sec, const vec3::Vector3Array &sp,
                                                                      const std::vector<double> &sY,
                                                                      const std::vector<double> &sZ,
                                                                      PrecisionFactor pf, ComputeMethod m) const
{
    if (sp.size() != sY.size() || sp.size() != sZ.size()) throw std::logic_error("Array sizes do not match!");
    if (m != GPU) cpuOnly("CoilGroup::computeAllMutualInductanceArrangements with CPU_ST/CPU_MT");
    std::vector<double> M, F, T;
    groupArrangements(memberCoils, sec, sp, sY, sZ, pf, CCOIL_PAIR_M, M, F, T);
    return M;
}

std::vector<std::pair<vec3::Vector3, vec3::Vector3>>
CoilGroup::computeAllForceTorqueArrangements(const Coil &sec, const vec3::Vector3Array &sp,
                                             const std::vector<double> &sY, const std::vector<double> &sZ,
                                             PrecisionFactor pf, ComputeMethod m) const
{
    if (sp.size() != sY.size() || sp.size() != sZ.size()) throw std::logic_error("Array sizes do not match!");
    if (m != GPU) cpuOnly("CoilGroup::computeAllForceTorqueArrangements with CPU_ST/CPU_MT");
    std::vector<double> M, F, T;
    groupArrangements(memberCoils, sec, sp, sY, sZ, pf, CCOIL_PAIR_FT, M, F, T);
    std::vector<std::pair<vec3::Vector3, vec3::Vector3>> out;
    out.reserve(sp.size());
    for (size_t i = 0; i < sp.size(); ++i)
        out.emplace_back(vec3::Vector3(F[3 * i], F[3 * i + 1], F[3 * i + 2]),
                         vec3::Vector3(T[3 * i], T[3 * i + 1], T[3 * i + 2]));
    return out;
}

CoilGroup::operator std::string() const
{
    std::ostringstream o;
    o << "CoilGroup(member_coils=[";
    for (size_t i = 0; i < memberCoils.size(); ++i) o << (i ? ", " : "") << std::string(*memberCoils[i]);
    o << "], default_precision_factor=" << std::string(defaultPrecisionFactor) << ", thread_count=" << threadCount
      << ")";
    return o.str();
}
