// CoilGroup of the C-Coil C++ API (interface of src/CoilGroup/CoilGroup.h:19-283). ComputeMethod::GPU
// evaluates all member coils in one ccoil_fields / ccoil_pairs call; CPU_ST / CPU_MT and the single-point
// methods belong to the reference's CPU path and throw std::logic_error.
#pragma once

#include "Coil.h"

#include <memory>
#include <string>
#include <vector>

class CoilGroup {
    std::vector<std::shared_ptr<Coil>> memberCoils;
    PrecisionFactor defaultPrecisionFactor;
    int threadCount{};

  public:
    explicit CoilGroup(std::vector<std::shared_ptr<Coil>> memberCoils = std::vector<std::shared_ptr<Coil>>(),
                       PrecisionFactor precisionFactor = PrecisionFactor(), int threadCount = g_defaultThreadCount);

    [[nodiscard]] PrecisionFactor getDefaultPrecisionFactor() const;
    [[nodiscard]] int getThreadCount() const;
    [[nodiscard]] const std::vector<std::shared_ptr<Coil>> &getMemberCoils() const;

    void setDefaultPrecisionFactor(PrecisionFactor precisionFactor = PrecisionFactor());
    void setThreadCount(int threadCount);
    void addCoil(double innerRadius, double thickness, double length, int numOfTurns, double current = 1.0,
                 PrecisionFactor precisionFactor = PrecisionFactor(), int coilThreads = g_defaultThreadCount,
                 vec3::Vector3 coordinatePosition = vec3::Vector3(), double yAxisAngle = 0.0, double zAxisAngle = 0.0);
    void removeCoil(size_t index);
    Coil &operator[](size_t index) const;
    [[nodiscard]] bool isPointInside(vec3::Vector3 pointVector);

    [[nodiscard]] vec3::Vector3 computeAPotentialVector(vec3::Vector3 pointVector) const;
    [[nodiscard]] vec3::Vector3 computeBFieldVector(vec3::Vector3 pointVector) const;
    [[nodiscard]] vec3::Vector3 computeEFieldVector(vec3::Vector3 pointVector) const;
    [[nodiscard]] vec3::Matrix3 computeBGradientMatrix(vec3::Vector3 pointVector) const;

    [[nodiscard]] vec3::Vector3Array computeAllAPotentialVectors(const vec3::Vector3Array &pointVectors,
                                                                 ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Vector3Array computeAllBFieldVectors(const vec3::Vector3Array &pointVectors,
                                                             ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Vector3Array computeAllEFieldVectors(const vec3::Vector3Array &pointVectors,
                                                             ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] vec3::Matrix3Array computeAllBGradientMatrices(const vec3::Vector3Array &pointVectors,
                                                                 ComputeMethod computeMethod = CPU_ST) const;

    [[nodiscard]] double computeMutualInductance(const Coil &secondary, PrecisionFactor precisionFactor = PrecisionFactor(),
                                                 ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] std::pair<vec3::Vector3, vec3::Vector3>
    computeForceTorque(const Coil &secondary, PrecisionFactor precisionFactor = PrecisionFactor(),
                       ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] std::pair<vec3::Vector3, vec3::Vector3> computeForceOnDipoleMoment(vec3::Vector3 pointVector,
                                                                                     vec3::Vector3 dipoleMoment) const;
    // Zero-copy GPU fields of all members at the group's precision factor (see Coil::computeAllFieldsGPU).
    void computeAllFieldsGPU(const double *points, size_t n, double *A, double *B, double *E, double *G,
                             bool devicePointers = false) const;
    // Force and torque on point dipoles in the group's field (GPU, true Jacobian; see Coil).
    [[nodiscard]] std::vector<std::pair<vec3::Vector3, vec3::Vector3>>
    computeAllForceOnDipoleMoments(const vec3::Vector3Array &points, const vec3::Vector3Array &moments,
                                   ComputeMethod computeMethod = GPU) const;

    [[nodiscard]] std::vector<double>
    computeAllMutualInductanceArrangements(const Coil &secondary, const vec3::Vector3Array &secondaryPositions,
                                           const std::vector<double> &secondaryYAngles,
                                           const std::vector<double> &secondaryZAngles,
                                           PrecisionFactor precisionFactor = PrecisionFactor(),
                                           ComputeMethod computeMethod = CPU_ST) const;
    [[nodiscard]] std::vector<std::pair<vec3::Vector3, vec3::Vector3>>
    computeAllForceTorqueArrangements(const Coil &secondary, const vec3::Vector3Array &secondaryPositions,
                                      const std::vector<double> &secondaryYAngles,
                                      const std::vector<double> &secondaryZAngles,
                                      PrecisionFactor precisionFactor = PrecisionFactor(),
                                      ComputeMethod computeMethod = CPU_ST) const;

    explicit operator std::string() const;
};
