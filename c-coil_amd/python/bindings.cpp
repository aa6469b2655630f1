// pybind11 module _c_coil (imported as c_coil through c_coil.py): submodules tensor, coil, coil_group with the reference's Python names
// (src/PyMain.cxx:36-43, src/Tensor/PyBindings.cxx, src/Coil/PyBindings.cxx:11-364,
// src/CoilGroup/PyBindings.cxx:13-113). The benchmark and compare submodules of the reference are not
// part of this GPU library.
#include "Coil.h"
#include "CoilGroup.h"
#include "ccoil_hip.h"

#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

namespace py = pybind11;

static void init_tensor(py::module_ &m)
{
    py::module_ t = m.def_submodule("tensor");
    py::class_<vec3::Vector3>(t, "Vector3")
        .def(py::init<>())
        .def(py::init<double, double, double>(), py::arg("x"), py::arg("y"), py::arg("z"))
        .def_readwrite("x", &vec3::Vector3::x)
        .def_readwrite("y", &vec3::Vector3::y)
        .def_readwrite("z", &vec3::Vector3::z)
        .def("__add__", &vec3::Vector3::operator+)
        .def("__iadd__", &vec3::Vector3::operator+=)
        .def("__sub__", &vec3::Vector3::operator-)
        .def("__isub__", &vec3::Vector3::operator-=)
        .def("__mul__", &vec3::Vector3::operator*)
        .def("__imul__", &vec3::Vector3::operator*=)
        .def("abs", &vec3::Vector3::abs)
        .def_static("scalar_product", &vec3::Vector3::scalarProduct, py::arg("vector1"), py::arg("vector2"))
        .def_static("cross_product", &vec3::Vector3::crossProduct, py::arg("vector1"), py::arg("vector2"))
        .def_static("get_from_cylindrical_coords", &vec3::Vector3::getFromCylindricalCoords, py::arg("z"),
                    py::arg("r"), py::arg("phi"))
        .def_static("get_from_spherical_coords", &vec3::Vector3::getFromSphericalCoords, py::arg("r"),
                    py::arg("theta"), py::arg("phi"))
        .def("get_as_cylindrical_coords", &vec3::Vector3::getAsCylindricalCoords)
        .def("get_as_spherical_coords", &vec3::Vector3::getAsSphericalCoords)
        .def("__repr__", &vec3::Vector3::operator std::string);

    py::class_<vec3::Matrix3>(t, "Matrix3")
        .def(py::init<>())
        .def(py::init<double, double, double, double, double, double, double, double, double>(),
             py::arg("xx") = 0.0, py::arg("xy") = 0.0, py::arg("xz") = 0.0, py::arg("yx") = 0.0, py::arg("yy") = 0.0,
             py::arg("yz") = 0.0, py::arg("zx") = 0.0, py::arg("zy") = 0.0, py::arg("zz") = 0.0)
        .def_readwrite("xx", &vec3::Matrix3::xx).def_readwrite("xy", &vec3::Matrix3::xy)
        .def_readwrite("xz", &vec3::Matrix3::xz).def_readwrite("yx", &vec3::Matrix3::yx)
        .def_readwrite("yy", &vec3::Matrix3::yy).def_readwrite("yz", &vec3::Matrix3::yz)
        .def_readwrite("zx", &vec3::Matrix3::zx).def_readwrite("zy", &vec3::Matrix3::zy)
        .def_readwrite("zz", &vec3::Matrix3::zz)
        .def("det", &vec3::Matrix3::det)
        .def("__add__", &vec3::Matrix3::operator+)
        .def("__iadd__", &vec3::Matrix3::operator+=)
        .def("__mul__", py::overload_cast<double>(&vec3::Matrix3::operator*, py::const_))
        .def("__mul__", py::overload_cast<const vec3::Matrix3 &>(&vec3::Matrix3::operator*, py::const_))
        .def("__mul__", py::overload_cast<const vec3::Vector3 &>(&vec3::Matrix3::operator*, py::const_))
        .def("__repr__", &vec3::Matrix3::operator std::string);

    py::class_<vec3::Triplet>(t, "Triplet")
        .def(py::init<>())
        .def(py::init<double, double, double>(), py::arg("first"), py::arg("second"), py::arg("third"))
        .def_readwrite("first", &vec3::Triplet::first)
        .def_readwrite("second", &vec3::Triplet::second)
        .def_readwrite("third", &vec3::Triplet::third)
        .def("__repr__", &vec3::Triplet::operator std::string);

    py::class_<vec3::Vector3Array>(t, "Vector3Array")
        .def(py::init<>())
        .def(py::init<size_t>(), py::arg("init_size"))
        .def(py::init<const std::vector<vec3::Vector3> &>(), py::arg("vector_list"))
        .def("append", py::overload_cast<const vec3::Vector3 &>(&vec3::Vector3Array::append), py::arg("appended_vector3"))
        .def("append", py::overload_cast<double, double, double>(&vec3::Vector3Array::append), py::arg("x"),
             py::arg("y"), py::arg("z"))
        .def("reserve", &vec3::Vector3Array::reserve, py::arg("reserve_size"))
        .def("resize", &vec3::Vector3Array::resize, py::arg("new_size"))
        .def("clear", &vec3::Vector3Array::clear)
        .def("size", &vec3::Vector3Array::size)
        .def("items", [](vec3::Vector3Array &a) { return a.getItems(); })
        .def("x", &vec3::Vector3Array::x)
        .def("y", &vec3::Vector3Array::y)
        .def("z", &vec3::Vector3Array::z)
        .def("abs", &vec3::Vector3Array::abs)
        .def("__len__", &vec3::Vector3Array::size)
        .def("__getitem__",
             [](vec3::Vector3Array &a, long long i) -> vec3::Vector3 & {
                 if (i < 0) i += (long long)a.size();
                 if (i < 0 || i >= (long long)a.size()) throw py::index_error("Array index out of range!");
                 return a[(size_t)i];
             },
             py::return_value_policy::reference_internal)
        .def("__setitem__",
             [](vec3::Vector3Array &a, long long i, const vec3::Vector3 &v) {
                 if (i < 0) i += (long long)a.size();
                 if (i < 0 || i >= (long long)a.size()) throw py::index_error("Array index out of range!");
                 a[(size_t)i] = v;
             })
        .def("__iadd__", &vec3::Vector3Array::operator+=)
        .def("__repr__", &vec3::Vector3Array::operator std::string);

    py::class_<vec3::Matrix3Array>(t, "Matrix3Array")
        .def(py::init<>())
        .def(py::init<size_t>(), py::arg("init_size"))
        .def(py::init<const std::vector<vec3::Matrix3> &>(), py::arg("matrix_list"))
        .def("append", py::overload_cast<const vec3::Matrix3 &>(&vec3::Matrix3Array::append), py::arg("appended_matrix3"))
        .def("reserve", &vec3::Matrix3Array::reserve, py::arg("reserve_size"))
        .def("resize", &vec3::Matrix3Array::resize, py::arg("new_size"))
        .def("clear", &vec3::Matrix3Array::clear)
        .def("size", &vec3::Matrix3Array::size)
        .def("items", [](vec3::Matrix3Array &a) { return a.getItems(); })
        .def("xx", &vec3::Matrix3Array::xx).def("xy", &vec3::Matrix3Array::xy).def("xz", &vec3::Matrix3Array::xz)
        .def("yx", &vec3::Matrix3Array::yx).def("yy", &vec3::Matrix3Array::yy).def("yz", &vec3::Matrix3Array::yz)
        .def("zx", &vec3::Matrix3Array::zx).def("zy", &vec3::Matrix3Array::zy).def("zz", &vec3::Matrix3Array::zz)
        .def("det", &vec3::Matrix3Array::det)
        .def("__len__", &vec3::Matrix3Array::size)
        .def("__getitem__",
             [](vec3::Matrix3Array &a, long long i) -> vec3::Matrix3 & {
                 if (i < 0) i += (long long)a.size();
                 if (i < 0 || i >= (long long)a.size()) throw py::index_error("Array index out of range!");
                 return a[(size_t)i];
             },
             py::return_value_policy::reference_internal)
        .def("__iadd__", &vec3::Matrix3Array::operator+=)
        .def("__repr__", &vec3::Matrix3Array::operator std::string);
}

static void init_coil(py::module_ &m)
{
    py::module_ c = m.def_submodule("coil");
    py::enum_<ComputeMethod>(c, "ComputeMethod")
        .value("CPU_ST", CPU_ST)
        .value("CPU_MT", CPU_MT)
        .value("GPU", GPU)
        .export_values();
    py::enum_<CoilType>(c, "CoilType")
        .value("RECTANGULAR", CoilType::RECTANGULAR)
        .value("THIN", CoilType::THIN)
        .value("FLAT", CoilType::FLAT)
        .value("FILAMENT", CoilType::FILAMENT)
        .export_values();

    py::class_<PrecisionFactor>(c, "PrecisionFactor")
        .def(py::init<>())
        .def(py::init<double>(), py::arg("relative_precision"))
        .def_readwrite("relative_precision", &PrecisionFactor::relativePrecision)
        .def("__repr__", &PrecisionFactor::operator std::string);

    py::class_<PrecisionArguments>(c, "PrecisionArguments")
        .def(py::init<>())
        .def(py::init<int, int, int, int, int, int>(), py::arg("angular_blocks"), py::arg("thickness_blocks"),
             py::arg("length_blocks"), py::arg("angular_increments"), py::arg("thickness_increments"),
             py::arg("length_increments"))
        .def_readwrite("angular_block_count", &PrecisionArguments::angularBlocks)
        .def_readwrite("thickness_block_count", &PrecisionArguments::thicknessBlocks)
        .def_readwrite("length_block_count", &PrecisionArguments::lengthBlocks)
        .def_readwrite("angular_increment_count", &PrecisionArguments::angularIncrements)
        .def_readwrite("thickness_increment_count", &PrecisionArguments::thicknessIncrements)
        .def_readwrite("length_increment_count", &PrecisionArguments::lengthIncrements)
        .def_static("get_coil_precision_arguments_CPU", &PrecisionArguments::getCoilPrecisionArgumentsCPU,
                    py::arg("coil"), py::arg("precision_factor"))
        .def_static("get_coil_precision_arguments_GPU", &PrecisionArguments::getCoilPrecisionArgumentsGPU,
                    py::arg("coil"), py::arg("precision_factor"))
        .def_static("get_secondary_coil_precision_arguments_GPU",
                    &PrecisionArguments::getSecondaryCoilPrecisionArgumentsGPU, py::arg("coil"),
                    py::arg("precision_factor"))
        .def("__repr__", &PrecisionArguments::operator std::string);

    py::class_<CoilPairArguments>(c, "CoilPairArguments")
        .def(py::init<>())
        .def(py::init<PrecisionArguments, PrecisionArguments>(), py::arg("primary_precision"),
             py::arg("secondary_precision"))
        .def_readwrite("primary_precision", &CoilPairArguments::primaryPrecision)
        .def_readwrite("secondary_precision", &CoilPairArguments::secondaryPrecision)
        .def_static("get_appropriate_coil_pair_arguments", &CoilPairArguments::getAppropriateCoilPairArguments,
                    py::arg("primary"), py::arg("secondary"), py::arg("precision_factor"),
                    py::arg("computeMethod") = CPU_ST, py::arg("z_axis_case") = false, py::arg("pure_GPU") = false)
        .def("__repr__", &CoilPairArguments::operator std::string);

    using V = vec3::Vector3;
    py::class_<Coil, std::shared_ptr<Coil>> coil(c, "Coil");
    coil.def(py::init<double, double, double, int, double, double, double, PrecisionFactor, int, V, double, double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("current"), py::arg("wire_resistivity"), py::arg("sine_frequency"),
             py::arg("precision_factor") = PrecisionFactor(), py::arg("thread_count") = g_defaultThreadCount,
             py::arg("coordinate_position") = V(), py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def(py::init<double, double, double, int, double, double, double, PrecisionArguments, PrecisionArguments,
                      int, V, double, double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("current"), py::arg("wire_resistivity"), py::arg("sine_frequency"),
             py::arg("precision_settings_CPU"), py::arg("precision_settings_GPU"),
             py::arg("thread_count") = g_defaultThreadCount, py::arg("coordinate_position") = V(),
             py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def(py::init<double, double, double, int, double, double, PrecisionFactor, int, V, double, double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("current"), py::arg("sine_frequency"), py::arg("precision_factor") = PrecisionFactor(),
             py::arg("thread_count") = g_defaultThreadCount, py::arg("coordinate_position") = V(),
             py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def(py::init<double, double, double, int, double, double, PrecisionArguments, PrecisionArguments, int, V,
                      double, double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("current"), py::arg("sine_frequency"), py::arg("precision_settings_CPU"),
             py::arg("precision_settings_GPU"), py::arg("thread_count") = g_defaultThreadCount,
             py::arg("coordinate_position") = V(), py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def(py::init<double, double, double, int, double, PrecisionFactor, int, V, double, double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("current"), py::arg("precision_factor") = PrecisionFactor(),
             py::arg("thread_count") = g_defaultThreadCount, py::arg("coordinate_position") = V(),
             py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def(py::init<double, double, double, int, double, PrecisionArguments, PrecisionArguments, int, V, double,
                      double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("current"), py::arg("precision_settings_CPU"), py::arg("precision_settings_GPU"),
             py::arg("thread_count") = g_defaultThreadCount, py::arg("coordinate_position") = V(),
             py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def(py::init<double, double, double, int, PrecisionFactor, int, V, double, double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("precision_factor") = PrecisionFactor(), py::arg("thread_count") = g_defaultThreadCount,
             py::arg("coordinate_position") = V(), py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def(py::init<double, double, double, int, PrecisionArguments, PrecisionArguments, int, V, double, double>(),
             py::arg("inner_radius"), py::arg("thickness"), py::arg("length"), py::arg("num_of_turns"),
             py::arg("precision_settings_CPU"), py::arg("precision_settings_GPU"),
             py::arg("thread_count") = g_defaultThreadCount, py::arg("coordinate_position") = V(),
             py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0);

    coil.def("get_id", &Coil::getId)
        .def("get_inner_radius", &Coil::getInnerRadius)
        .def("get_thickness", &Coil::getThickness)
        .def("get_length", &Coil::getLength)
        .def("get_num_of_turns", &Coil::getNumOfTurns)
        .def("get_current_density", &Coil::getCurrentDensity)
        .def("get_current", &Coil::getCurrent)
        .def("get_wire_resistivity", &Coil::getWireResistivity)
        .def("is_sine_driven", &Coil::isSineDriven)
        .def("get_sine_frequency", &Coil::getSineFrequency)
        .def("get_magnetic_moment", &Coil::getMagneticMoment)
        .def("get_average_wire_thickness", &Coil::getAverageWireThickness)
        .def("get_self_inductance", &Coil::getSelfInductance)
        .def("get_resistance", &Coil::getResistance)
        .def("get_reactance", &Coil::getReactance)
        .def("get_impedance", &Coil::getImpedance)
        .def("get_precision_settings_CPU", &Coil::getPrecisionSettingsCPU)
        .def("get_precision_settings_GPU", &Coil::getPrecisionSettingsGPU)
        .def("get_thread_count", &Coil::getThreadCount)
        .def("is_using_fast_method", &Coil::isUsingFastMethod)
        .def("get_coil_type", &Coil::getCoilType)
        .def("get_position_vector", &Coil::getPositionVector)
        .def("get_rotation_angles", &Coil::getRotationAngles)
        .def("set_current_density", &Coil::setCurrentDensity, py::arg("current_density"))
        .def("set_current", &Coil::setCurrent, py::arg("current"))
        .def("set_wire_resistivity", &Coil::setWireResistivity, py::arg("wire_resistivity"))
        .def("set_sine_frequency", &Coil::setSineFrequency, py::arg("sine_frequency"))
        .def("set_default_precision_GPU",
             static_cast<void (Coil::*)(const PrecisionArguments &)>(&Coil::setDefaultPrecisionGPU),
             py::arg("precision_settings"))
        .def("set_default_precision_GPU", static_cast<void (Coil::*)(PrecisionFactor)>(&Coil::setDefaultPrecisionGPU),
             py::arg("precision_factor") = PrecisionFactor())
        .def("set_default_precision_CPU",
             static_cast<void (Coil::*)(const PrecisionArguments &)>(&Coil::setDefaultPrecisionCPU),
             py::arg("precision_settings"))
        .def("set_default_precision_CPU", static_cast<void (Coil::*)(PrecisionFactor)>(&Coil::setDefaultPrecisionCPU),
             py::arg("precision_factor") = PrecisionFactor())
        .def("set_thread_count", &Coil::setThreadCount, py::arg("thread_count"))
        .def("set_self_inductance", &Coil::setSelfInductance, py::arg("self_inductance"))
        .def("set_position_and_orientation", &Coil::setPositionAndOrientation, py::arg("position_vector"),
             py::arg("y_axis_angle"), py::arg("z_axis_angle"))
        .def("is_point_inside", &Coil::isPointInside, py::arg("point_vector"));

    using VA = vec3::Vector3Array;
    using MA = vec3::Matrix3Array;
    coil.def("compute_B_field_vector", py::overload_cast<V>(&Coil::computeBFieldVector, py::const_), py::arg("point_vector"))
        .def("compute_B_field_vector", py::overload_cast<V, const PrecisionArguments &>(&Coil::computeBFieldVector, py::const_),
             py::arg("point_vector"), py::arg("used_precision"))
        .def("compute_A_potential_vector", py::overload_cast<V>(&Coil::computeAPotentialVector, py::const_),
             py::arg("point_vector"))
        .def("compute_A_potential_vector",
             py::overload_cast<V, const PrecisionArguments &>(&Coil::computeAPotentialVector, py::const_),
             py::arg("point_vector"), py::arg("used_precision"))
        .def("compute_E_field_vector", py::overload_cast<V>(&Coil::computeEFieldVector, py::const_), py::arg("point_vector"))
        .def("compute_E_field_vector", py::overload_cast<V, const PrecisionArguments &>(&Coil::computeEFieldVector, py::const_),
             py::arg("point_vector"), py::arg("used_precision"))
        .def("compute_B_gradient_matrix", py::overload_cast<V>(&Coil::computeBGradientMatrix, py::const_),
             py::arg("point_vector"))
        .def("compute_B_gradient_matrix",
             py::overload_cast<V, const PrecisionArguments &>(&Coil::computeBGradientMatrix, py::const_),
             py::arg("point_vector"), py::arg("used_precision"))
        .def("compute_all_B_field_vectors", py::overload_cast<const VA &, ComputeMethod>(&Coil::computeAllBFieldVectors, py::const_),
             py::arg("point_vectors"), py::arg("compute_method") = CPU_ST)
        .def("compute_all_B_field_vectors",
             py::overload_cast<const VA &, const PrecisionArguments &, ComputeMethod>(&Coil::computeAllBFieldVectors, py::const_),
             py::arg("point_vectors"), py::arg("used_precision"), py::arg("compute_method") = CPU_ST)
        .def("compute_all_A_potential_vectors",
             py::overload_cast<const VA &, ComputeMethod>(&Coil::computeAllAPotentialVectors, py::const_),
             py::arg("point_vectors"), py::arg("compute_method") = CPU_ST)
        .def("compute_all_A_potential_vectors",
             py::overload_cast<const VA &, const PrecisionArguments &, ComputeMethod>(&Coil::computeAllAPotentialVectors, py::const_),
             py::arg("point_vectors"), py::arg("used_precision"), py::arg("compute_method") = CPU_ST)
        .def("compute_all_E_field_vectors", py::overload_cast<const VA &, ComputeMethod>(&Coil::computeAllEFieldVectors, py::const_),
             py::arg("point_vectors"), py::arg("compute_method") = CPU_ST)
        .def("compute_all_E_field_vectors",
             py::overload_cast<const VA &, const PrecisionArguments &, ComputeMethod>(&Coil::computeAllEFieldVectors, py::const_),
             py::arg("point_vectors"), py::arg("used_precision"), py::arg("compute_method") = CPU_ST)
        .def("compute_all_B_gradient_matrices",
             py::overload_cast<const VA &, ComputeMethod>(&Coil::computeAllBGradientMatrices, py::const_),
             py::arg("point_vectors"), py::arg("compute_method") = CPU_ST)
        .def("compute_all_B_gradient_matrices",
             py::overload_cast<const VA &, const PrecisionArguments &, ComputeMethod>(&Coil::computeAllBGradientMatrices, py::const_),
             py::arg("point_vectors"), py::arg("used_precision"), py::arg("compute_method") = CPU_ST)
        .def_static("compute_mutual_inductance",
                    py::overload_cast<const Coil &, const Coil &, PrecisionFactor, ComputeMethod>(&Coil::computeMutualInductance),
                    py::arg("primary"), py::arg("secondary"), py::arg("precision_factor") = PrecisionFactor(),
                    py::arg("compute_method") = CPU_ST)
        .def_static("compute_mutual_inductance",
                    py::overload_cast<const Coil &, const Coil &, const CoilPairArguments &, ComputeMethod>(
                        &Coil::computeMutualInductance),
                    py::arg("primary"), py::arg("secondary"), py::arg("inductance_arguments"),
                    py::arg("compute_method") = CPU_ST)
        .def("compute_secondary_induced_voltage",
             py::overload_cast<const Coil &, PrecisionFactor, ComputeMethod>(&Coil::computeSecondaryInducedVoltage, py::const_),
             py::arg("secondary"), py::arg("precision_factor") = PrecisionFactor(), py::arg("compute_method") = CPU_ST)
        .def("compute_secondary_induced_voltage",
             py::overload_cast<const Coil &, const CoilPairArguments &, ComputeMethod>(&Coil::computeSecondaryInducedVoltage,
                                                                                    py::const_),
             py::arg("secondary"), py::arg("inductance_arguments"), py::arg("compute_method") = CPU_ST)
        .def("compute_and_set_self_inductance", &Coil::computeAndSetSelfInductance,
             py::arg("precision_factor") = PrecisionFactor())
        .def_static("compute_ampere_force",
                    py::overload_cast<const Coil &, const Coil &, PrecisionFactor, ComputeMethod>(&Coil::computeForceTorque),
                    py::arg("primary"), py::arg("secondary"), py::arg("precision_factor") = PrecisionFactor(),
                    py::arg("compute_method") = CPU_ST)
        .def_static("compute_ampere_force",
                    py::overload_cast<const Coil &, const Coil &, const CoilPairArguments &, ComputeMethod>(
                        &Coil::computeForceTorque),
                    py::arg("primary"), py::arg("secondary"), py::arg("force_arguments"),
                    py::arg("compute_method") = CPU_ST)
        .def("compute_force_on_dipole_moment",
             py::overload_cast<V, V>(&Coil::computeForceOnDipoleMoment, py::const_), py::arg("point_vector"),
             py::arg("dipole_moment"))
        .def("compute_force_on_dipole_moment",
             py::overload_cast<V, V, const PrecisionArguments &>(&Coil::computeForceOnDipoleMoment, py::const_),
             py::arg("point_vector"), py::arg("dipole_moment"), py::arg("used_precision"))
        .def("compute_all_force_on_dipole_moments", &Coil::computeAllForceOnDipoleMoments, py::arg("point_vectors"),
             py::arg("dipole_moments"), py::arg("compute_method") = GPU)
        .def_static("compute_all_mutual_inductance_arrangements", &Coil::computeAllMutualInductanceArrangements,
                    py::arg("primary"), py::arg("secondary"), py::arg("primary_positions"),
                    py::arg("secondary_positions"), py::arg("primary_Y_angles"), py::arg("primary_Z_angles"),
                    py::arg("secondary_Y_angles"), py::arg("secondary_Z_angles"),
                    py::arg("precision_factor") = PrecisionFactor(), py::arg("compute_method") = CPU_ST)
        .def_static("compute_all_force_torque_arrangements", &Coil::computeAllForceTorqueArrangements,
                    py::arg("primary"), py::arg("secondary"), py::arg("primary_positions"),
                    py::arg("secondary_positions"), py::arg("primary_Y_angles"), py::arg("primary_Z_angles"),
                    py::arg("secondary_Y_angles"), py::arg("secondary_Z_angles"),
                    py::arg("precision_factor") = PrecisionFactor(), py::arg("compute_method") = CPU_ST)
        .def_static("set_gpu_precision", &Coil::setGpuPrecision, py::arg("precision"))
        .def_static("get_gpu_precision", &Coil::getGpuPrecision)
        .def("__repr__", &Coil::operator std::string);

    c.attr("GPU_FP64") = (int)CCOIL_FP64;
    c.attr("GPU_FP32C") = (int)CCOIL_FP32C;
    c.attr("GPU_FP32") = (int)CCOIL_FP32;
    c.attr("GPU_FP64_REFERENCE") = (int)CCOIL_FP64_REFERENCE;
    c.attr("GPU_FP64_IEEE") = (int)CCOIL_FP64_IEEE;
    c.def("set_gpu_increments", &ccoil::setGpuIncrements, py::arg("increments"));
    c.def("get_gpu_increments", &ccoil::gpuIncrements);
}

static void init_coil_group(py::module_ &m)
{
    py::module_ g = m.def_submodule("coil_group");
    py::class_<CoilGroup>(g, "CoilGroup")
        .def(py::init<std::vector<std::shared_ptr<Coil>>, PrecisionFactor, int>(),
             py::arg("member_coils") = std::vector<std::shared_ptr<Coil>>(),
             py::arg("precision_factor") = PrecisionFactor(), py::arg("thread_count") = g_defaultThreadCount)
        .def("get_default_precision_factor", &CoilGroup::getDefaultPrecisionFactor)
        .def("get_thread_count", &CoilGroup::getThreadCount)
        .def("get_member_coils", &CoilGroup::getMemberCoils)
        .def("set_default_precision_factor", &CoilGroup::setDefaultPrecisionFactor,
             py::arg("precision_factor") = PrecisionFactor())
        .def("set_thread_count", &CoilGroup::setThreadCount, py::arg("thread_count"))
        .def("add_coil", &CoilGroup::addCoil, py::arg("inner_radius"), py::arg("thickness"), py::arg("length"),
             py::arg("num_of_turns"), py::arg("current") = 1.0, py::arg("precision_factor") = PrecisionFactor(),
             py::arg("coil_threads") = g_defaultThreadCount, py::arg("coordinate_position") = vec3::Vector3(),
             py::arg("y_axis_angle") = 0.0, py::arg("z_axis_angle") = 0.0)
        .def("remove_coil", &CoilGroup::removeCoil, py::arg("index"))
        .def("__getitem__",
             [](const CoilGroup &self, long long i) -> std::shared_ptr<Coil> {
                 const long long n = (long long)self.getMemberCoils().size();
                 if (i < 0) i += n;
                 if (i < 0 || i >= n) throw py::index_error("Array index out of range!");
                 return self.getMemberCoils()[(size_t)i];
             })
        .def("__len__", [](const CoilGroup &self) { return self.getMemberCoils().size(); })
        .def("is_point_inside", &CoilGroup::isPointInside, py::arg("point_vector"))
        .def("compute_B_field_vector", &CoilGroup::computeBFieldVector, py::arg("point_vector"))
        .def("compute_A_potential_vector", &CoilGroup::computeAPotentialVector, py::arg("point_vector"))
        .def("compute_E_field_vector", &CoilGroup::computeEFieldVector, py::arg("point_vector"))
        .def("compute_B_gradient_matrix", &CoilGroup::computeBGradientMatrix, py::arg("point_vector"))
        .def("compute_all_A_potential_vectors", &CoilGroup::computeAllAPotentialVectors, py::arg("point_vectors"),
             py::arg("compute_method") = CPU_ST)
        .def("compute_all_B_field_vectors", &CoilGroup::computeAllBFieldVectors, py::arg("point_vectors"),
             py::arg("compute_method") = CPU_ST)
        .def("compute_all_E_field_vectors", &CoilGroup::computeAllEFieldVectors, py::arg("point_vectors"),
             py::arg("compute_method") = CPU_ST)
        .def("compute_all_B_gradient_matrices", &CoilGroup::computeAllBGradientMatrices, py::arg("point_vectors"),
             py::arg("compute_method") = CPU_ST)
        .def("compute_mutual_inductance", &CoilGroup::computeMutualInductance, py::arg("secondary"),
             py::arg("precision_factor") = PrecisionFactor(), py::arg("compute_method") = CPU_ST)
        .def("compute_force_torque", &CoilGroup::computeForceTorque, py::arg("secondary"),
             py::arg("precision_factor") = PrecisionFactor(), py::arg("compute_method") = CPU_ST)
        .def("compute_force_on_dipole_moment", &CoilGroup::computeForceOnDipoleMoment, py::arg("point_vector"),
             py::arg("dipole_moment"))
        .def("compute_all_force_on_dipole_moments", &CoilGroup::computeAllForceOnDipoleMoments,
             py::arg("point_vectors"), py::arg("dipole_moments"), py::arg("compute_method") = GPU)
        .def("compute_all_mutual_inductance_arrangements", &CoilGroup::computeAllMutualInductanceArrangements,
             py::arg("secondary"), py::arg("secondary_positions"), py::arg("secondary_y_angles"),
             py::arg("secondary_z_angles"), py::arg("precision_factor") = PrecisionFactor(),
             py::arg("compute_method") = CPU_ST)
        .def("compute_all_force_torque_arrangements", &CoilGroup::computeAllForceTorqueArrangements,
             py::arg("secondary"), py::arg("secondary_positions"), py::arg("secondary_y_angles"),
             py::arg("secondary_z_angles"), py::arg("precision_factor") = PrecisionFactor(),
             py::arg("compute_method") = CPU_ST)
        .def("__repr__", &CoilGroup::operator std::string);
}

PYBIND11_MODULE(_c_coil, m)
{
    m.doc() = "C-Coil (MI355X build): circular-coil magnetostatics on gfx950 through the ccoil_hip C ABI";
    init_tensor(m);
    init_coil(m);
    init_coil_group(m);
    m.attr("__version__") = ccoil_version();
}
