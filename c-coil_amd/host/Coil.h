// C-Coil C++ API, MI355X build: Coil, PrecisionFactor, PrecisionArguments, CoilPairArguments.
//
// The public interface mirrors src/Coil/Coil.h:30-814 of the reference so user code compiles unchanged.
// ComputeMethod::GPU runs on the HIP path through the C ABI (include/ccoil_hip.h). CPU_ST / CPU_MT and the
// single-point methods are the reference's CPU implementation and are not part of this GPU library: they
// throw std::logic_error naming the GPU entry point to use instead.
#pragma once

#include "Tensor.h"

#include <string>
#include <utility>
#include <vector>

enum ComputeMethod { CPU_ST, CPU_MT, GPU };                  // src/Coil/EnumsAndConstants/ComputeMethod.h:4
enum class CoilType { RECTANGULAR, THIN, FLAT, FILAMENT };  // src/Coil/EnumsAndConstants/CoilType.h:5

const int g_defaultThreadCount = 8;  // Coil.h:18

namespace ccoil {
// Quadrature-order cap of the GPU planning (the reference's compile-time GPU_INCREMENTS, CMakeLists.txt:16;
// here a runtime setting in [1, 100], default 80, also read from the CCOIL_GPU_INCREMENTS environment variable).
int gpuIncrements();
void setGpuIncrements(int increments);
}  // namespace ccoil

class Coil;

struct PrecisionFactor {
    PrecisionFactor();
    explicit PrecisionFactor(double relativePrecision);
    double relativePrecision;
    explicit operator std::string() const;
};

struct PrecisionArguments {
    PrecisionArguments();
    explicit PrecisionArguments(int angularBlocks, int thicknessBlocks, int lengthBlocks, int angularIncrements,
                                int thicknessIncrements, int lengthIncrements);
    int angularBlocks, thicknessBlocks, lengthBlocks;
    int angularIncrements, thicknessIncrements, lengthIncrements;

    static PrecisionArguments getCoilPrecisionArgumentsCPU(const Coil &coil, PrecisionFactor precisionFactor);
    static PrecisionArguments getCoilPrecisionArgumentsGPU(const Coil &coil, PrecisionFactor precisionFactor);
    static PrecisionArguments getSecondaryCoilPrecisionArgumentsGPU(const Coil &coil, PrecisionFactor precisionFactor);
    explicit operator std::string() const;

  private:
    static PrecisionArguments calculatePrecisionArguments(const Coil &coil, PrecisionFactor precisionFactor,
                                                          bool useGPU);
};

struct CoilPairArguments {
    CoilPairArguments();
    explicit CoilPairArguments(const PrecisionArguments &primaryPrecision, const PrecisionArguments &secondaryPrecision);
    PrecisionArguments primaryPrecision;
    PrecisionArguments secondaryPrecision;

    static CoilPairArguments getAppropriateCoilPairArguments(const Coil &primary, const Coil &secondary,
                                                             PrecisionFactor precisionFactor,
                                                             ComputeMethod computeMethod = CPU_ST,
                                                             bool zAxisCase = false, bool pureGPU = false);
    explicit operator std::string() const;

    static std::vector<std::pair<int, int>> balanceIncrements(int totalIncrements,
                                                              const std::vector<std::pair<int, double>> &components);
};

class CoilGroup;

class Coil {
    friend CoilGroup;

    unsigned long long id;
    double innerRadius, thickness, length;
    int numOfTurns;
    double currentDensity{}, current{};
    double wireResistivity{};
    bool sineDriven{};
    double sineFrequency{};
    double magneticMoment{}, averageWireThickness{};
    double resistance{}, selfInductance{}, reactance{}, impedance{};
    CoilType coilType{};
    bool useFastMethod{};
    int threadCount{};
    PrecisionArguments defaultPrecisionCPU, defaultPrecisionGPU;
    vec3::Vector3 positionVector{};
    double yAxisAngle{}, zAxisAngle{};
    vec3::Matrix3 transformationMatrix{}, inverseTransformationMatrix{};

  public:
    Coil(double innerRadius, double thickness, double length, int numOfTurns, double current, double wireResistivity,
         double sineFrequency, PrecisionFactor precisionFactor = PrecisionFactor(),
         int threadCount = g_defaultThreadCount, vec3::Vector3 coordinatePosition = vec3::Vector3(),
         double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    Coil(double innerRadius, double thickness, double length, int numOfTurns, double current, double wireResistivity,
         double sineFrequency, const PrecisionArguments &precisionSettingsCPU,
         const PrecisionArguments &precisionSettingsGPU, int threadCount = g_defaultThreadCount,
         vec3::Vector3 coordinatePosition = vec3::Vector3(), double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    Coil(double innerRadius, double thickness, double length, int numOfTurns, double current, double sineFrequency,
         PrecisionFactor precisionFactor = PrecisionFactor(), int threadCount = g_defaultThreadCount,
         vec3::Vector3 coordinatePosition = vec3::Vector3(), double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    Coil(double innerRadius, double thickness, double length, int numOfTurns, double current, double sineFrequency,
         const PrecisionArguments &precisionSettingsCPU, const PrecisionArguments &precisionSettingsGPU,
         int threadCount = g_defaultThreadCount, vec3::Vector3 coordinatePosition = vec3::Vector3(),
         double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    Coil(double innerRadius, double thickness, double length, int numOfTurns, double current,
         PrecisionFactor precisionFactor = PrecisionFactor(), int threadCount = g_defaultThreadCount,
         vec3::Vector3 coordinatePosition = vec3::Vector3(), double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    Coil(double innerRadius, double thickness, double length, int numOfTurns, double current,
         const PrecisionArguments &precisionSettingsCPU, const PrecisionArguments &precisionSettingsGPU,
         int threadCount = g_defaultThreadCount, vec3::Vector3 coordinatePosition = vec3::Vector3(),
         double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    Coil(double innerRadius, double thickness, double length, int numOfTurns,
         PrecisionFactor precisionFactor = PrecisionFactor(), int threadCount = g_defaultThreadCount,
         vec3::Vector3 coordinatePosition = vec3::Vector3(), double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    Coil(double innerRadius, double thickness, double length, int numOfTurns,
         const PrecisionArguments &precisionSettingsCPU, const PrecisionArguments &precisionSettingsGPU,
         int threadCount = g_defaultThreadCount, vec3::Vector3 coordinatePosition = vec3::Vector3(),
         double yAxisAngle = 0.0, double zAxisAngle = 0.0);

    [[nodiscard]] unsigned long long getId() const;
    [[nodiscard]] double getInnerRadius() const;
    [[nodiscard]] double getThickness() const;
    [[nodiscard]] double getLength() const;
    [[nodiscard]] int getNumOfTurns() const;
    [[nodiscard]] double getCurrentDensity() const;
    [[nodiscard]] double getCurrent() const;
    [[nodiscard]] double getWireResistivity() const;
    [[nodiscard]] bool isSineDriven() const;
    [[nodiscard]] double getSineFrequency() const;
    [[nodiscard]] vec3::Vector3 getMagneticMoment();
    [[nodiscard]] double getAverageWireThickness() const;
    [[nodiscard]] double getSelfInductance() const;
    [[nodiscard]] double getResistance();
    [[nodiscard]] double getReactance();
    [[nodiscard]] double getImpedance();
    [[nodiscard]] const PrecisionArguments &getPrecisionSettingsCPU() const;
    [[nodiscard]] const PrecisionArguments &getPrecisionSettingsGPU() const;
    [[nodiscard]] int getThreadCount() const;
    [[nodiscard]] bool isUsingFastMethod() const;
    [[nodiscard]] CoilType getCoilType() const;
    [[nodiscard]] vec3::Vector3 getPositionVector() const;
    [[nodiscard]] std::pair<double, double> getRotationAngles() const;
    [[nodiscard]] vec3::Matrix3 getTransformationMatrix() const;
    [[nodiscard]] vec3::Matrix3 getInverseTransformationMatrix() const;

    void setCurrentDensity(double currentDensity);
    void setCurrent(double current);
    void setWireResistivity(double wireResistivity);
    void setSineFrequency(double sineFrequency);
    void setDefaultPrecisionCPU(const PrecisionArguments &precisionSettings);
    void setDefaultPrecisionCPU(PrecisionFactor precisionFactor = PrecisionFactor());
    void setDefaultPrecisionGPU(const PrecisionArguments &precisionSettings);
    void setDefaultPrecisionGPU(PrecisionFactor precisionFactor = PrecisionFactor());
    void setDefaultPrecision(PrecisionFactor precisionFactor = PrecisionFactor());
    void setThreadCount(int threadCount);
    void setPositionAndOrientation(vec3::Vector3 positionVector = vec3::Vector3(), double yAxisAngle = 0.0,
                                   double zAxisAngle = 0.0);
    void setSelfInductance(double selfInductance);
    [[nodiscard]] bool isPointInside(vec3::Vector3 pointVector);

    // Single-point evaluation is the reference's CPU path: throws std::logic_error here.
    [[nodiscard]] vec3::Vector3 computeAPotentialVector(vec3::Vector3 pointVector) const;
    [[nodiscard]] vec3::Vector3 computeAPotentialVector(vec3::Vector3 pointVector, const PrecisionArguments &p) const;
    [[nodiscard]] vec3::Vector3 computeBFieldVector(vec3::Vector3 pointVector) const;
    [[nodiscard]] vec3::Vector3 computeBFieldVector(vec3::Vector3 pointVector, const PrecisionArguments &p) const;
    [[nodiscard]] vec3::Vector3 computeEFieldVector(vec3::Vector3 pointVector) const;
    [[nodiscard]] vec3::Vector3 computeEFieldVector(vec3::Vector3 pointVector, const PrecisionArguments &p) const;
    [[nodiscard]] vec3::Matrix3 computeBGradientMatrix(vec3::Vector3 pointVector) const;
    [[nodiscard]] vec3::Matrix3 computeBGradientMatrix(vec3::Vector3 pointVector, const PrecisionArguments &p) const;

    [[nodiscard]] vec3::Vector3Array computeAllAPotentialVectors(const vec3::Vector3Array &pointVectors,
                                                                 ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Vector3Array computeAllAPotentialVectors(const vec3::Vector3Array &pointVectors,
                                                                 const PrecisionArguments &usedPrecision,
                                                                 ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Vector3Array computeAllBFieldVectors(const vec3::Vector3Array &pointVectors,
                                                             ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Vector3Array computeAllBFieldVectors(const vec3::Vector3Array &pointVectors,
                                                             const PrecisionArguments &usedPrecision,
                                                             ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Vector3Array computeAllEFieldVectors(const vec3::Vector3Array &pointVectors,
                                                             ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Vector3Array computeAllEFieldVectors(const vec3::Vector3Array &pointVectors,
                                                             const PrecisionArguments &usedPrecision,
                                                             ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Matrix3Array computeAllBGradientMatrices(const vec3::Vector3Array &pointVectors,
                                                                 ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Matrix3Array computeAllBGradientMatrices(const vec3::Vector3Array &pointVectors,
                                                                 const PrecisionArguments &usedPrecision,
                                                                 ComputeMethod computeMethod = CPU_ST) const;

    static double computeMutualInductance(const Coil &primary, const Coil &secondary,
                                          PrecisionFactor precisionFactor = PrecisionFactor(),
                                          ComputeMethod computeMethod = CPU_ST);
    static double computeMutualInductance(const Coil &primary, const Coil &secondary,
                                          const CoilPairArguments &inductanceArguments,
                                          ComputeMethod computeMethod = CPU_ST);
    [[nodiscard]] double computeSecondaryInducedVoltage(const Coil &secondary,
                                                        PrecisionFactor precisionFactor = PrecisionFactor(),
                                                        ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] double computeSecondaryInducedVoltage(const Coil &secondary,
                                                        const CoilPairArguments &inductanceArguments,
                                                        ComputeMethod computeMethod = CPU_ST) const;
    // Self-inductance by the reference's coaxial rule (CalculateMInductanceZAxis.cxx:224-317), evaluated on the
    // device for THIN and RECTANGULAR coils (ccoil_self_inductance) and through the device field path for FLAT
    // coils; FILAMENT coils have no self-inductance and throw std::logic_error, as in the reference.
    double computeAndSetSelfInductance(PrecisionFactor precisionFactor);

    static std::pair<vec3::Vector3, vec3::Vector3>
    computeForceTorque(const Coil &primary, const Coil &secondary, PrecisionFactor precisionFactor = PrecisionFactor(),
                       ComputeMethod computeMethod = CPU_ST);
    static std::pair<vec3::Vector3, vec3::Vector3>
    computeForceTorque(const Coil &primary, const Coil &secondary, const CoilPairArguments &forceArguments,
                       ComputeMethod computeMethod = CPU_ST);
    // Single-point (CPU) in the reference: throws std::logic_error here.
    [[nodiscard]] std::pair<vec3::Vector3, vec3::Vector3> computeForceOnDipoleMoment(vec3::Vector3 pointVector,
                                                                                     vec3::Vector3 dipoleMoment) const;
    [[nodiscard]] std::pair<vec3::Vector3, vec3::Vector3>
    computeForceOnDipoleMoment(vec3::Vector3 pointVector, vec3::Vector3 dipoleMoment,
                               const PrecisionArguments &usedPrecision) const;

    // Force and torque on point dipoles (moments[k] at points[k]) in this coil's field, GPU only: F = J^T m with
    // the true Jacobian of B (the reference's single-point form multiplies its T*G gradient, ComputeForceMethods.cxx:38-49).
    [[nodiscard]] std::vector<std::pair<vec3::Vector3, vec3::Vector3>>
    computeAllForceOnDipoleMoments(const vec3::Vector3Array &points, const vec3::Vector3Array &moments,
                                   ComputeMethod computeMethod = GPU) const;

    // Zero-copy GPU fields at the default GPU precision (no counterpart in the reference; backs the numpy and
    // DLPack entry points of the Python module): points are n AoS xyz doubles and every non-null output is written in
    // place (A, B, E: 3 doubles per point; G: 9, the reference's T*G_local), all in one launch. devicePointers: all
    // buffers are device pointers on the context's device, ordered on its stream.
    void computeAllFieldsGPU(const double *points, size_t n, double *A, double *B, double *E, double *G,
                             bool devicePointers = false) const;

    static std::vector<double> computeAllMutualInductanceArrangements(
        const Coil &primary, const Coil &secondary, const vec3::Vector3Array &primaryPositions,
        const vec3::Vector3Array &secondaryPositions, const std::vector<double> &primaryYAngles,
        const std::vector<double> &primaryZAngles, const std::vector<double> &secondaryYAngles,
        const std::vector<double> &secondaryZAngles, PrecisionFactor precisionFactor = PrecisionFactor(),
        ComputeMethod computeMethod = CPU_ST);
    static std::vector<std::pair<vec3::Vector3, vec3::Vector3>> computeAllForceTorqueArrangements(
        const Coil &primary, const Coil &secondary, const vec3::Vector3Array &primaryPositions,
        const vec3::Vector3Array &secondaryPositions, const std::vector<double> &primaryYAngles,
        const std::vector<double> &primaryZAngles, const std::vector<double> &secondaryYAngles,
        const std::vector<double> &secondaryZAngles, PrecisionFactor precisionFactor = PrecisionFactor(),
        ComputeMethod computeMethod = CPU_ST);

    explicit operator std::string() const;

    // Arithmetic of the GPU quadrature for this thread (CCOIL_FP64, CCOIL_FP32C, CCOIL_FP32 of ccoil_hip.h).
    // The reference's TYPE build option (CMakeLists.txt:18) as a runtime setting; default FP32C.
    static void setGpuPrecision(int precision);
    static int getGpuPrecision();
    // The far-field series of GPU field calls for this thread (no counterpart in the reference): -1 the precision's
    // default (on in FP32C, off in FP64), 0 off (every pair by the rule), 1 on.
    static void setGpuFarField(int mode);
    static int getGpuFarField();

    static bool isZAxisCase(const Coil &primary, const Coil &secondary);
    static std::pair<bool, double> improvedPrecisionCase(const Coil &primary, const Coil &secondary);

  private:
    void calculateMagneticMoment();
    void calculateAverageWireThickness();
    void calculateResistance();
    void calculateReactance();
    void calculateImpedance();
    void calculateCoilType();
    void calculateTransformationMatrices();
};
