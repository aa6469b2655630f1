// Coil: construction, attributes and the GPU compute paths.
// Attribute semantics follow src/Coil/Coil.cxx:22-264 and src/Coil/Utils/CalculateAttributes.cxx:12-78;
// GPU semantics follow the reference's CUDA path (src/Coil/Fields/CalculateAllFieldsGPU.cxx,
// src/Coil/MInductance/*, src/Coil/ForceAndTorque/*), evaluated by the HIP kernels behind ccoil_hip.h.
#include "Coil.h"
#include "GpuBridge.h"

#include <atomic>
#include <cmath>
#include <sstream>
#include <stdexcept>

using ccoil::bridge::CoilSoA;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMuReduced = 1e-7;
constexpr double kThinRatio = 1e-7;         // g_thinCoilApproximationRatio
constexpr double kZAxisRatio = 1e-12;       // g_zAxisApproximationRatio
constexpr double kCommonPlaneRatio = 1e-14; // g_commonPlaneApproximationRatio
constexpr double kDefaultCurrent = 1.0, kDefaultResistivity = 1.63e-8, kDefaultFrequency = 50.0;

std::atomic<unsigned long long> g_next_id{0};

[[noreturn]] void cpuOnly(const char *what)
{
    throw std::logic_error(std::string(what) +
                           " is computed on the CPU by C-Coil; this library provides the GPU path "
                           "(use the computeAll* / arrangement methods with ComputeMethod::GPU)");
}

void requireGpu(ComputeMethod m, const char *what)
{
    if (m != GPU) cpuOnly(what);
}

double clampThin(double v, double R) { return v / R < kThinRatio ? 1e-18 * R * kThinRatio : v; }

}  // namespace

Coil::Coil(double R, double a, double b, int N, double I, double rho, double f, const PrecisionArguments &cpu,
           const PrecisionArguments &gpu, int threads, vec3::Vector3 pos, double yA, double zA)
    : id(++g_next_id), innerRadius(R), thickness(clampThin(a, R)), length(clampThin(b, R)), numOfTurns(N),
      defaultPrecisionCPU(cpu), defaultPrecisionGPU(gpu), positionVector(pos), yAxisAngle(yA), zAxisAngle(zA)
{
    useFastMethod = b / R >= kThinRatio;
    calculateTransformationMatrices();
    calculateCoilType();
    setCurrent(I);
    calculateAverageWireThickness();
    setWireResistivity(rho);
    setSineFrequency(f);
    setThreadCount(threads);
}

Coil::Coil(double R, double a, double b, int N, double I, double rho, double f, PrecisionFactor pf, int threads,
           vec3::Vector3 pos, double yA, double zA)
    : Coil(R, a, b, N, I, rho, f, PrecisionArguments(), PrecisionArguments(), threads, pos, yA, zA)
{
    setDefaultPrecision(pf);
}

Coil::Coil(double R, double a, double b, int N, double I, double f, PrecisionFactor pf, int threads,
           vec3::Vector3 pos, double yA, double zA)
    : Coil(R, a, b, N, I, kDefaultResistivity, f, pf, threads, pos, yA, zA) {}

Coil::Coil(double R, double a, double b, int N, double I, double f, const PrecisionArguments &cpu,
           const PrecisionArguments &gpu, int threads, vec3::Vector3 pos, double yA, double zA)
    : Coil(R, a, b, N, I, kDefaultResistivity, f, cpu, gpu, threads, pos, yA, zA) {}

Coil::Coil(double R, double a, double b, int N, double I, PrecisionFactor pf, int threads, vec3::Vector3 pos,
           double yA, double zA)
    : Coil(R, a, b, N, I, kDefaultResistivity, kDefaultFrequency, pf, threads, pos, yA, zA) {}

Coil::Coil(double R, double a, double b, int N, double I, const PrecisionArguments &cpu,
           const PrecisionArguments &gpu, int threads, vec3::Vector3 pos, double yA, double zA)
    : Coil(R, a, b, N, I, kDefaultResistivity, kDefaultFrequency, cpu, gpu, threads, pos, yA, zA) {}

Coil::Coil(double R, double a, double b, int N, PrecisionFactor pf, int threads, vec3::Vector3 pos, double yA,
           double zA)
    : Coil(R, a, b, N, kDefaultCurrent, kDefaultResistivity, kDefaultFrequency, pf, threads, pos, yA, zA) {}

Coil::Coil(double R, double a, double b, int N, const PrecisionArguments &cpu, const PrecisionArguments &gpu,
           int threads, vec3::Vector3 pos, double yA, double zA)
    : Coil(R, a, b, N, kDefaultCurrent, kDefaultResistivity, kDefaultFrequency, cpu, gpu, threads, pos, yA, zA) {}

unsigned long long Coil::getId() const { return id; }
double Coil::getInnerRadius() const { return innerRadius; }
double Coil::getThickness() const { return thickness; }
double Coil::getLength() const { return length; }
int Coil::getNumOfTurns() const { return numOfTurns; }
double Coil::getCurrentDensity() const { return currentDensity; }
double Coil::getCurrent() const { return current; }
double Coil::getWireResistivity() const { return wireResistivity; }
bool Coil::isSineDriven() const { return sineDriven; }
double Coil::getSineFrequency() const { return sineFrequency; }
double Coil::getAverageWireThickness() const { return averageWireThickness; }
double Coil::getSelfInductance() const { return selfInductance; }
const PrecisionArguments &Coil::getPrecisionSettingsCPU() const { return defaultPrecisionCPU; }
const PrecisionArguments &Coil::getPrecisionSettingsGPU() const { return defaultPrecisionGPU; }
int Coil::getThreadCount() const { return threadCount; }
bool Coil::isUsingFastMethod() const { return useFastMethod; }
CoilType Coil::getCoilType() const { return coilType; }
vec3::Vector3 Coil::getPositionVector() const { return positionVector; }
std::pair<double, double> Coil::getRotationAngles() const { return {yAxisAngle, zAxisAngle}; }
vec3::Matrix3 Coil::getTransformationMatrix() const { return transformationMatrix; }
vec3::Matrix3 Coil::getInverseTransformationMatrix() const { return inverseTransformationMatrix; }

vec3::Vector3 Coil::getMagneticMoment()
{
    calculateMagneticMoment();
    const double sA = std::sin(yAxisAngle), cA = std::cos(yAxisAngle);
    const double sB = std::sin(zAxisAngle), cB = std::cos(zAxisAngle);
    return vec3::Vector3(magneticMoment * sA * cB, magneticMoment * sA * sB, magneticMoment * cA);
}

double Coil::getResistance() { calculateResistance(); return resistance; }
double Coil::getReactance() { calculateReactance(); return reactance; }
double Coil::getImpedance() { calculateImpedance(); return impedance; }

void Coil::setCurrentDensity(double j)
{
    currentDensity = j;
    current = j * length * thickness / numOfTurns;
}

void Coil::setCurrent(double i)
{
    current = i;
    currentDensity = i * numOfTurns / (length * thickness);
}

void Coil::setWireResistivity(double r) { wireResistivity = r; }

void Coil::setSineFrequency(double f)
{
    sineDriven = f > 0.0;
    sineFrequency = sineDriven ? f : 0.0;
    calculateImpedance();
}

void Coil::setDefaultPrecisionCPU(const PrecisionArguments &p) { defaultPrecisionCPU = p; }
void Coil::setDefaultPrecisionCPU(PrecisionFactor pf)
{
    defaultPrecisionCPU = PrecisionArguments::getCoilPrecisionArgumentsCPU(*this, pf);
}
void Coil::setDefaultPrecisionGPU(const PrecisionArguments &p) { defaultPrecisionGPU = p; }
void Coil::setDefaultPrecisionGPU(PrecisionFactor pf)
{
    defaultPrecisionGPU = PrecisionArguments::getCoilPrecisionArgumentsGPU(*this, pf);
}
void Coil::setDefaultPrecision(PrecisionFactor pf)
{
    setDefaultPrecisionCPU(pf);
    setDefaultPrecisionGPU(pf);
}
void Coil::setThreadCount(int t) { threadCount = t; }

void Coil::setPositionAndOrientation(vec3::Vector3 p, double yA, double zA)
{
    positionVector = p;
    yAxisAngle = yA;
    zAxisAngle = zA;
    calculateTransformationMatrices();
}

void Coil::setSelfInductance(double L) { selfInductance = L; }

bool Coil::isPointInside(vec3::Vector3 p)
{
    // Coil.cxx:250-261
    const vec3::Vector3 l = inverseTransformationMatrix * (p - positionVector);
    const vec3::Triplet cyl = l.getAsCylindricalCoords();
    return std::abs(cyl.first) < 0.5 * length + kThinRatio * innerRadius &&
           std::abs(cyl.second - innerRadius - 0.5 * thickness) < 0.5 * thickness + kThinRatio * innerRadius;
}

void Coil::calculateMagneticMoment()
{
    magneticMoment = kPi * current * numOfTurns *
                     (innerRadius * innerRadius + innerRadius * thickness + thickness * thickness / 3);
}

void Coil::calculateAverageWireThickness() { averageWireThickness = std::sqrt(length * thickness / numOfTurns); }

void Coil::calculateResistance()
{
    // CalculateAttributes.cxx:21-34 (skin-effect corrected DC resistance)
    const double wr = averageWireThickness * 0.5;
    const double ohmic = wireResistivity * numOfTurns * 2 * kPi * (innerRadius + thickness * 0.5) / (wr * wr * kPi);
    const double skin = std::sqrt(wireResistivity / (kPi * sineFrequency * kMuReduced));
    const double ohmicSurface = kPi * wr * wr;
    const double sineSurface = 2 * kPi * (skin * skin * (std::exp(-wr / skin) - 1) + skin * wr);
    resistance = ohmic * (ohmicSurface / sineSurface);
}

void Coil::calculateReactance() { reactance = selfInductance * 2 * kPi * sineFrequency; }

void Coil::calculateImpedance()
{
    calculateResistance();
    calculateReactance();
    impedance = std::sqrt(resistance * resistance + reactance * reactance);
}

void Coil::calculateCoilType()
{
    const bool thinA = thickness / innerRadius < kThinRatio, thinB = length / innerRadius < kThinRatio;
    coilType = (thinA && thinB) ? CoilType::FILAMENT : thinA ? CoilType::THIN : thinB ? CoilType::FLAT
                                                                                      : CoilType::RECTANGULAR;
}

void Coil::calculateTransformationMatrices()
{
    // CalculateAttributes.cxx:63-78: T = Rz(beta) Ry(alpha) Rz(-beta), T^-1 = T^T
    const double cY = std::cos(yAxisAngle), sY = std::sin(yAxisAngle);
    const double cZ = std::cos(zAxisAngle), sZ = std::sin(zAxisAngle);
    transformationMatrix = vec3::Matrix3(cZ * cZ * cY - sZ * sZ, -sZ * cZ * cY - sZ * cZ, cZ * sY,
                                         sZ * cZ * cY + sZ * cZ, cZ * cZ - sZ * sZ * cY, sZ * sY,
                                         -sY * cZ, sY * sZ, cY);
    inverseTransformationMatrix = vec3::Matrix3(cZ * cZ * cY - sZ * sZ, sZ * cZ * cY + sZ * cZ, -cZ * sY,
                                                -sZ * cZ * cY - sZ * cZ, cZ * cZ - sZ * sZ * cY, sZ * sY,
                                                sY * cZ, sY * sZ, cY);
}

// --------------------------------------------------------------------------------------------
// Single-point (CPU) methods.
// --------------------------------------------------------------------------------------------
vec3::Vector3 Coil::computeAPotentialVector(vec3::Vector3) const { cpuOnly("Coil::computeAPotentialVector"); }
vec3::Vector3 Coil::computeAPotentialVector(vec3::Vector3, const PrecisionArguments &) const
{
    cpuOnly("Coil::computeAPotentialVector");
}
vec3::Vector3 Coil::computeBFieldVector(vec3::Vector3) const { cpuOnly("Coil::computeBFieldVector"); }
vec3::Vector3 Coil::computeBFieldVector(vec3::Vector3, const PrecisionArguments &) const
{
    cpuOnly("Coil::computeBFieldVector");
}
vec3::Vector3 Coil::computeEFieldVector(vec3::Vector3) const { cpuOnly("Coil::computeEFieldVector"); }
vec3::Vector3 Coil::computeEFieldVector(vec3::Vector3, const PrecisionArguments &) const
{
    cpuOnly("Coil::computeEFieldVector");
}
vec3::Matrix3 Coil::computeBGradientMatrix(vec3::Vector3) const { cpuOnly("Coil::computeBGradientMatrix"); }
vec3::Matrix3 Coil::computeBGradientMatrix(vec3::Vector3, const PrecisionArguments &) const
{
    cpuOnly("Coil::computeBGradientMatrix");
}
std::pair<vec3::Vector3, vec3::Vector3> Coil::computeForceOnDipoleMoment(vec3::Vector3, vec3::Vector3) const
{
    cpuOnly("Coil::computeForceOnDipoleMoment");
}
std::pair<vec3::Vector3, vec3::Vector3> Coil::computeForceOnDipoleMoment(vec3::Vector3, vec3::Vector3,
                                                                         const PrecisionArguments &) const
{
    cpuOnly("Coil::computeForceOnDipoleMoment");
}
// ComputeMInductanceMethods.cxx:55-80: the coil is moved to the origin with zero angles so that the coaxial rules
// apply, with the pair arguments the reference plans for (coil, coil) on the CPU. THIN/RECTANGULAR coils take
// calculateSelfInductance; FLAT coils take the slow coaxial mutual-inductance rule with itself, whose field is the
// device field rule (the primary's increments in one block).
double Coil::computeAndSetSelfInductance(PrecisionFactor pf)
{
    if (coilType == CoilType::FILAMENT)
        throw std::logic_error("Coil loop self inductance is not defined, integral is divergent");
    const vec3::Vector3 position = positionVector;
    const double yA = yAxisAngle, zA = zAxisAngle;
    setPositionAndOrientation();
    double L = 0.0;
    try {
        const auto args = CoilPairArguments::getAppropriateCoilPairArguments(*this, *this, pf, CPU_ST, true, false);
        L = coilType == CoilType::FLAT ? ccoil::bridge::zAxisMutualInductance(*this, *this, args)
                                       : ccoil::bridge::selfInductance(*this, args);
    } catch (...) {
        setPositionAndOrientation(position, yA, zA);
        throw;
    }
    setSelfInductance(L);
    setPositionAndOrientation(position, yA, zA);
    return L;
}

// --------------------------------------------------------------------------------------------
// GPU field paths (CalculateAllFieldsGPU.cxx:16-309 -> ccoil_fields with a one-coil descriptor).
// --------------------------------------------------------------------------------------------
namespace {

void fieldsGPU(const Coil &c, const PrecisionArguments &args, const vec3::Vector3Array &pts, uint32_t q,
               double *out)
{
    CoilSoA s;
    s.add(c, args);
    const ccoil_coils v = s.view();
    ccoil_ctx ctx = ccoil::bridge::context();
    double *A = (q & CCOIL_A) ? out : nullptr, *B = (q & CCOIL_B) ? out : nullptr;
    double *E = (q & CCOIL_E) ? out : nullptr, *G = (q & CCOIL_G) ? out : nullptr;
    ccoil::bridge::check(ccoil_fields(ctx, &v, pts.data(), (int64_t)pts.size(), q, ccoil::bridge::precision(), 0, A,
                                      B, E, G, nullptr),
                         ctx, "ccoil_fields");
}

}  // namespace

vec3::Vector3Array Coil::computeAllAPotentialVectors(const vec3::Vector3Array &pts, const PrecisionArguments &p,
                                                     ComputeMethod m) const
{
    requireGpu(m, "Coil::computeAllAPotentialVectors with CPU_ST/CPU_MT");
    vec3::Vector3Array out(pts.size());
    fieldsGPU(*this, p, pts, CCOIL_A, out.data());
    return out;
}

vec3::Vector3Array Coil::computeAllAPotentialVectors(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    return computeAllAPotentialVectors(pts, m == GPU ? defaultPrecisionGPU : defaultPrecisionCPU, m);
}

vec3::Vector3Array Coil::computeAllBFieldVectors(const vec3::Vector3Array &pts, const PrecisionArguments &p,
                                                 ComputeMethod m) const
{
    requireGpu(m, "Coil::computeAllBFieldVectors with CPU_ST/CPU_MT");
    vec3::Vector3Array out(pts.size());
    fieldsGPU(*this, p, pts, CCOIL_B, out.data());
    return out;
}

vec3::Vector3Array Coil::computeAllBFieldVectors(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    return computeAllBFieldVectors(pts, m == GPU ? defaultPrecisionGPU : defaultPrecisionCPU, m);
}

vec3::Vector3Array Coil::computeAllEFieldVectors(const vec3::Vector3Array &pts, const PrecisionArguments &p,
                                                 ComputeMethod m) const
{
    requireGpu(m, "Coil::computeAllEFieldVectors with CPU_ST/CPU_MT");
    vec3::Vector3Array out(pts.size());
    fieldsGPU(*this, p, pts, CCOIL_E, out.data());
    return out;
}

vec3::Vector3Array Coil::computeAllEFieldVectors(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    return computeAllEFieldVectors(pts, m == GPU ? defaultPrecisionGPU : defaultPrecisionCPU, m);
}

vec3::Matrix3Array Coil::computeAllBGradientMatrices(const vec3::Vector3Array &pts, const PrecisionArguments &p,
                                                     ComputeMethod m) const
{
    requireGpu(m, "Coil::computeAllBGradientMatrices with CPU_ST/CPU_MT");
    vec3::Matrix3Array out(pts.size());
    fieldsGPU(*this, p, pts, CCOIL_G, out.data());
    return out;
}

vec3::Matrix3Array Coil::computeAllBGradientMatrices(const vec3::Vector3Array &pts, ComputeMethod m) const
{
    return computeAllBGradientMatrices(pts, m == GPU ? defaultPrecisionGPU : defaultPrecisionCPU, m);
}

std::vector<std::pair<vec3::Vector3, vec3::Vector3>>
Coil::computeAllForceOnDipoleMoments(const vec3::Vector3Array &points, const vec3::Vector3Array &moments,
                                     ComputeMethod m) const
{
    requireGpu(m, "Coil::computeAllForceOnDipoleMoments with CPU_ST/CPU_MT");
    CoilSoA s;
    s.add(*this, defaultPrecisionGPU);
    return ccoil::bridge::dipoleForceTorque(s, points, moments);
}

void Coil::computeAllFieldsGPU(const double *points, size_t n, double *A, double *B, double *E, double *G,
                               bool devicePointers) const
{
    CoilSoA s;
    s.add(*this, defaultPrecisionGPU);
    ccoil::bridge::fieldsInto(s, points, n, A, B, E, G, devicePointers);
}

// --------------------------------------------------------------------------------------------
// Pairs.
// --------------------------------------------------------------------------------------------
bool Coil::isZAxisCase(const Coil &p, const Coil &s)
{
    // SupportFunctions.cxx:55-70
    const vec3::Vector3 a = p.positionVector, b = s.positionVector;
    return std::abs(a.x / p.innerRadius) < kZAxisRatio && std::abs(a.y / p.innerRadius) < kZAxisRatio &&
           std::abs(b.x / p.innerRadius) < kZAxisRatio && std::abs(b.y / p.innerRadius) < kZAxisRatio &&
           std::abs(p.yAxisAngle / (2 * kPi)) < kZAxisRatio && std::abs(s.yAxisAngle / (2 * kPi)) < kZAxisRatio;
}

std::pair<bool, double> Coil::improvedPrecisionCase(const Coil &prim, const Coil &sec)
{
    // SupportFunctions.cxx:124-161: coplanar axes allow integrating half of the secondary ring.
    const vec3::Vector3 pn(std::sin(prim.yAxisAngle) * std::cos(prim.zAxisAngle),
                           std::sin(prim.yAxisAngle) * std::sin(prim.zAxisAngle), std::cos(prim.yAxisAngle));
    const vec3::Vector3 sn(std::sin(sec.yAxisAngle) * std::cos(sec.zAxisAngle),
                           std::sin(sec.yAxisAngle) * std::sin(sec.zAxisAngle), std::cos(sec.yAxisAngle));
    const vec3::Vector3 diff = prim.positionVector - sec.positionVector;
    vec3::Vector3 plane = vec3::Vector3::crossProduct(pn, diff);
    const bool improved = std::abs(vec3::Vector3::scalarProduct(plane, sn)) <= kCommonPlaneRatio;
    double offset = 0.0;
    if (improved) {
        auto zero = [](const vec3::Vector3 &v) { return v.x == 0.0 && v.y == 0.0 && v.z == 0.0; };
        if (zero(plane)) plane = vec3::Vector3::crossProduct(sn, diff);
        if (zero(plane)) plane = vec3::Vector3::crossProduct(sn, pn);
        offset = std::atan2(std::cos(sec.yAxisAngle) * (plane.x * std::cos(sec.zAxisAngle) +
                                                        plane.y * std::sin(sec.zAxisAngle)) -
                                plane.z * std::sin(sec.yAxisAngle),
                            plane.x * std::sin(sec.zAxisAngle) - plane.y * std::cos(sec.zAxisAngle));
    }
    return {improved, offset};
}

namespace {

// One pair through ccoil_pairs with the general secondary rule (blocks, coplanar half ring):
// CalculateMInductanceGeneral.cxx:8-117, CalculateForceGeneral.cxx:8-131.
void pairGPU(const Coil &prim, const Coil &sec, const CoilPairArguments &args, uint32_t what, bool improvedForM,
             double *M, double *F, double *T)
{
    CoilSoA p, s;
    p.add(prim, args.primaryPrecision);
    s.add(sec, args.secondaryPrecision);
    const ccoil_coils pv = p.view(), sv = s.view();
    double offset = 0.0;
    int32_t half = 0;
    if (improvedForM) {
        auto ip = Coil::improvedPrecisionCase(prim, sec);
        half = ip.first ? 1 : 0;
        offset = ip.second;
    }
    ccoil_ctx ctx = ccoil::bridge::context();
    ccoil::bridge::check(ccoil_pairs(ctx, &pv, &sv, nullptr, nullptr, 1, &offset, &half, what,
                                     ccoil::bridge::precision(), 0, M, F, T, nullptr, nullptr),
                         ctx, "ccoil_pairs");
}

}  // namespace

// ComputeMInductanceMethods.cxx:7-41: coaxial pairs take the z-axis rules, every other pair the general rule.
double Coil::computeMutualInductance(const Coil &prim, const Coil &sec, const CoilPairArguments &args,
                                     ComputeMethod m)
{
    requireGpu(m, "Coil::computeMutualInductance with CPU_ST/CPU_MT");
    if (isZAxisCase(prim, sec)) return ccoil::bridge::zAxisMutualInductance(prim, sec, args);
    double M = 0.0;
    pairGPU(prim, sec, args, CCOIL_PAIR_M, true, &M, nullptr, nullptr);
    return M;
}

double Coil::computeMutualInductance(const Coil &prim, const Coil &sec, PrecisionFactor pf, ComputeMethod m)
{
    requireGpu(m, "Coil::computeMutualInductance with CPU_ST/CPU_MT");
    auto args = CoilPairArguments::getAppropriateCoilPairArguments(prim, sec, pf, m, isZAxisCase(prim, sec), false);
    return computeMutualInductance(prim, sec, args, m);
}

double Coil::computeSecondaryInducedVoltage(const Coil &sec, const CoilPairArguments &args, ComputeMethod m) const
{
    return computeMutualInductance(*this, sec, args, m) * 2 * kPi * sineFrequency;
}

double Coil::computeSecondaryInducedVoltage(const Coil &sec, PrecisionFactor pf, ComputeMethod m) const
{
    return computeMutualInductance(*this, sec, pf, m) * 2 * kPi * sineFrequency;
}

std::pair<vec3::Vector3, vec3::Vector3> Coil::computeForceTorque(const Coil &prim, const Coil &sec,
                                                                 const CoilPairArguments &args, ComputeMethod m)
{
    requireGpu(m, "Coil::computeForceTorque with CPU_ST/CPU_MT");
    // ComputeForceMethods.cxx:5-24: a coaxial pair has only the axial force
    if (isZAxisCase(prim, sec))
        return {vec3::Vector3(0.0, 0.0, ccoil::bridge::zAxisForce(prim, sec, args)), vec3::Vector3()};
    double F[3], T[3];
    pairGPU(prim, sec, args, CCOIL_PAIR_FT, false, nullptr, F, T);
    return {vec3::Vector3(F[0], F[1], F[2]), vec3::Vector3(T[0], T[1], T[2])};
}

std::pair<vec3::Vector3, vec3::Vector3> Coil::computeForceTorque(const Coil &prim, const Coil &sec,
                                                                 PrecisionFactor pf, ComputeMethod m)
{
    requireGpu(m, "Coil::computeForceTorque with CPU_ST/CPU_MT");
    auto args = CoilPairArguments::getAppropriateCoilPairArguments(prim, sec, pf, m, isZAxisCase(prim, sec), false);
    return computeForceTorque(prim, sec, args, m);
}

namespace {

void checkArrangementSizes(size_t n, const vec3::Vector3Array &sp, const std::vector<double> &a,
                           const std::vector<double> &b, const std::vector<double> &c, const std::vector<double> &d)
{
    if (n != sp.size() || n != a.size() || n != b.size() || n != c.size() || n != d.size())
        throw std::logic_error("Array sizes do not match!");
}

// Pure-GPU arrangement rule (CalculateMInductanceArrangements.cxx:84-153, MInductance.cu:24-157): one
// planned argument set, every configuration its own (primary pose, secondary pose) pair.
void arrangementsGPU(const Coil &prim, const Coil &sec, const vec3::Vector3Array &pp, const vec3::Vector3Array &sp,
                     const std::vector<double> &pY, const std::vector<double> &pZ, const std::vector<double> &sY,
                     const std::vector<double> &sZ, PrecisionFactor pf, uint32_t what, double *M, double *F,
                     double *T)
{
    const auto args = CoilPairArguments::getAppropriateCoilPairArguments(prim, sec, pf, GPU, false, true);
    CoilSoA p, s;
    for (size_t i = 0; i < pp.size(); ++i) {
        p.add(prim, args.primaryPrecision, pp[i], pY[i], pZ[i]);
        s.add(sec, args.secondaryPrecision, sp[i], sY[i], sZ[i]);
    }
    const ccoil_coils pv = p.view(), sv = s.view();
    ccoil_ctx ctx = ccoil::bridge::context();
    // The reference's single-pair force arrangement scales by 2 pi N2 / I1 (CalculateForceArrangements.cxx:125).
    const uint32_t flags = (what & CCOIL_PAIR_FT) ? CCOIL_FORCE_REFERENCE_SCALE : 0;
    ccoil::bridge::check(ccoil_pairs(ctx, &pv, &sv, nullptr, nullptr, (int64_t)pp.size(), nullptr, nullptr, what,
                                     ccoil::bridge::precision(), flags, M, F, T, nullptr, nullptr),
                         ctx, "ccoil_pairs");
}

}  // namespace

std::vector<double> Coil::computeAllMutualInductanceArrangements(
    const Coil &prim, const Coil &sec, const vec3::Vector3Array &pp, const vec3::Vector3Array &sp,
    const std::vector<double> &pY, const std::vector<double> &pZ, const std::vector<double> &sY,
    const std::vector<double> &sZ, PrecisionFactor pf, ComputeMethod m)
{
    checkArrangementSizes(pp.size(), sp, pY, pZ, sY, sZ);
    requireGpu(m, "Coil::computeAllMutualInductanceArrangements with CPU_ST/CPU_MT");
    std::vector<double> M(pp.size());
    if (!pp.size()) return M;
    arrangementsGPU(prim, sec, pp, sp, pY, pZ, sY, sZ, pf, CCOIL_PAIR_M, M.data(), nullptr, nullptr);
    return M;
}

std::vector<std::pair<vec3::Vector3, vec3::Vector3>> Coil::computeAllForceTorqueArrangements(
    const Coil &prim, const Coil &sec, const vec3::Vector3Array &pp, const vec3::Vector3Array &sp,
    const std::vector<double> &pY, const std::vector<double> &pZ, const std::vector<double> &sY,
    const std::vector<double> &sZ, PrecisionFactor pf, ComputeMethod m)
{
    checkArrangementSizes(pp.size(), sp, pY, pZ, sY, sZ);
    requireGpu(m, "Coil::computeAllForceTorqueArrangements with CPU_ST/CPU_MT");
    std::vector<std::pair<vec3::Vector3, vec3::Vector3>> out;
    if (!pp.size()) return out;
    std::vector<double> F(3 * pp.size()), T(3 * pp.size());
    arrangementsGPU(prim, sec, pp, sp, pY, pZ, sY, sZ, pf, CCOIL_PAIR_FT, nullptr, F.data(), T.data());
    out.reserve(pp.size());
    for (size_t i = 0; i < pp.size(); ++i)
        out.emplace_back(vec3::Vector3(F[3 * i], F[3 * i + 1], F[3 * i + 2]),
                         vec3::Vector3(T[3 * i], T[3 * i + 1], T[3 * i + 2]));
    return out;
}

void Coil::setGpuPrecision(int p) { ccoil::bridge::setPrecision(p); }
int Coil::getGpuPrecision() { return ccoil::bridge::precision(); }
void Coil::setGpuFarField(int mode) { ccoil::bridge::setFarField(mode); }
int Coil::getGpuFarField() { return ccoil::bridge::farField(); }

Coil::operator std::string() const
{
    std::ostringstream o;
    o << "Coil(id=" << id << ", inner_radius=" << innerRadius << ", thickness=" << thickness
      << ", length=" << length << ", num_of_turns=" << numOfTurns << ", current_density=" << currentDensity
      << ", current=" << current << ", wire_resistivity=" << wireResistivity << ", is_sine_driven=" << sineDriven
      << ", sine_frequency=" << sineFrequency << ", magnetic_moment=" << magneticMoment
      << ", average_wire_thickness=" << averageWireThickness << ", resistance=" << resistance
      << ", self_inductance=" << selfInductance << ", reactance=" << reactance << ", impedance=" << impedance
      << ", use_fast_method=" << useFastMethod << ", thread_count=" << threadCount
      << ", default_precision_CPU=" << std::string(defaultPrecisionCPU)
      << ", default_precision_GPU=" << std::string(defaultPrecisionGPU)
      << ", position_vector=" << std::string(positionVector) << ", y_axis_angle=" << yAxisAngle
      << ", z_axis_angle=" << zAxisAngle << ")";
    return o.str();
}
