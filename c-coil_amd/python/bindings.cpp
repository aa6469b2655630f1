// pybind11 module _c_coil (imported as c_coil through c_coil.py): submodules tensor, coil, coil_group with the reference's Python names
// (src/PyMain.cxx:36-43, src/Tensor/PyBindings.cxx, src/Coil/PyBindings.cxx:11-364,
// src/CoilGroup/PyBindings.cxx:13-113). The benchmark and compare submodules of the reference are not
// part of this GPU library.
#include "Coil.h"
#include "CoilGroup.h"
#include "GpuBridge.h"
#include "ccoil_hip.h"

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <stdexcept>

namespace py = pybind11;

// ------------------------------------------------------------------------------------------------------------------
// Zero-copy buffers: numpy arrays and DLPack producers (torch tensors on the CPU or on the GPU, ...). No counterpart
// in the reference, whose Python API converts every call through Vector3Array / Matrix3Array.
// ------------------------------------------------------------------------------------------------------------------
namespace zc {

// The DLPack C ABI (dlpack.h of the DLPack specification), legacy unversioned capsule "dltensor".
struct DLDevice {
    int32_t device_type;
    int32_t device_id;
};
struct DLDataType {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
};
struct DLTensor {
    void *data;
    DLDevice device;
    int32_t ndim;
    DLDataType dtype;
    int64_t *shape;
    int64_t *strides;
    uint64_t byte_offset;
};
struct DLManagedTensor {
    DLTensor dl_tensor;
    void *manager_ctx;
    void (*deleter)(DLManagedTensor *);
};
enum : int32_t { kDLCPU = 1, kDLCUDA = 2, kDLCUDAHost = 3, kDLROCM = 10, kDLROCMHost = 11 };
enum : uint8_t { kDLFloat = 2 };

struct Buffer {
    double *ptr = nullptr;
    int64_t rows = 0;
    bool device = false;
    py::object keep;   // the numpy array or the DLPack capsule: keeps the producer's memory alive for the call
};

// Rows of `width` doubles, row-major and contiguous: (n, 3) for points, A, B, E; (n, 9) or (n, 3, 3) for G.
bool shapeOk(int ndim, const int64_t *shape, int width)
{
    if (ndim == 2) return shape[1] == width;
    return ndim == 3 && width == 9 && shape[1] == 3 && shape[2] == 3;
}

Buffer view(py::handle obj, int width, bool output, const char *name)
{
    const std::string who(name);
    Buffer b;
    if (py::isinstance<py::array>(obj)) {
        py::array arr = py::reinterpret_borrow<py::array>(obj);
        if (output) {
            if (!arr.dtype().is(py::dtype::of<double>()) || !(arr.flags() & py::array::c_style) || !arr.writeable())
                throw std::invalid_argument(who + " must be a writable C-contiguous float64 array");
        } else {
            arr = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(obj);
            if (!arr) throw std::invalid_argument(who + " cannot be read as a float64 array");
        }
        std::vector<int64_t> shape(arr.shape(), arr.shape() + arr.ndim());
        if (!shapeOk((int)arr.ndim(), shape.data(), width))
            throw std::invalid_argument(who + (width == 3 ? " must have shape (n, 3)" : " must have shape (n, 9) or (n, 3, 3)"));
        b.ptr = static_cast<double *>(arr.mutable_data());
        b.rows = shape[0];
        b.keep = arr;
        return b;
    }
    if (!py::hasattr(obj, "__dlpack__") || !py::hasattr(obj, "__dlpack_device__"))
        throw std::invalid_argument(who + " must be a numpy array or a DLPack array (e.g. a torch tensor)");
    py::tuple dev = obj.attr("__dlpack_device__")();
    const int type = dev[0].cast<int>(), id = dev[1].cast<int>();
    py::object cap;
    if (type == kDLCUDA || type == kDLROCM) {
        if (id != ccoil::bridge::device())
            throw std::invalid_argument(who + " is on device " + std::to_string(id) + ", the context uses device " +
                                        std::to_string(ccoil::bridge::device()));
        // the producer orders its pending work before the context's stream
        cap = obj.attr("__dlpack__")(py::arg("stream") = (intptr_t)ccoil::bridge::stream());
        b.device = true;
    } else if (type == kDLCPU || type == kDLCUDAHost || type == kDLROCMHost) {
        cap = obj.attr("__dlpack__")();
    } else {
        throw std::invalid_argument(who + ": unsupported DLPack device type " + std::to_string(type));
    }
    auto *mt = static_cast<DLManagedTensor *>(PyCapsule_GetPointer(cap.ptr(), "dltensor"));
    if (!mt) throw py::error_already_set();
    const DLTensor &t = mt->dl_tensor;
    if (t.dtype.code != kDLFloat || t.dtype.bits != 64 || t.dtype.lanes != 1)
        throw std::invalid_argument(who + " must be float64");
    if (!shapeOk(t.ndim, t.shape, width))
        throw std::invalid_argument(who + (width == 3 ? " must have shape (n, 3)" : " must have shape (n, 9) or (n, 3, 3)"));
    if (t.strides) {
        int64_t expect = 1;
        for (int k = t.ndim - 1; k >= 0; --k) {
            if (t.shape[k] != 1 && t.strides[k] != expect) throw std::invalid_argument(who + " must be contiguous");
            expect *= t.shape[k];
        }
    }
    b.ptr = reinterpret_cast<double *>(static_cast<char *>(t.data) + t.byte_offset);
    b.rows = t.shape[0];
    b.keep = cap;   // the capsule's destructor returns the tensor to its producer
    return b;
}

[[noreturn]] void cpuOnly(const char *what)
{
    throw std::logic_error(std::string(what) + " with CPU_ST/CPU_MT is computed on the CPU by C-Coil; this library "
                                               "provides the GPU path (compute_method=GPU)");
}

// Fields into caller buffers: all of one side (host or device), one launch for every requested quantity.
template <class Source>
void into(const Source &src, py::handle points, py::handle A, py::handle B, py::handle E, py::handle G)
{
    Buffer p = view(points, 3, false, "points");
    Buffer out[4];
    const py::handle outs[4] = {A, B, E, G};
    const char *names[4] = {"A", "B", "E", "G"};
    bool any = false;
    for (int k = 0; k < 4; ++k) {
        if (outs[k].is_none()) continue;
        out[k] = view(outs[k], k == 3 ? 9 : 3, true, names[k]);
        if (out[k].rows != p.rows) throw std::invalid_argument(std::string(names[k]) + " must have as many rows as points");
        if (out[k].device != p.device) throw std::invalid_argument("points and outputs must all be host or all device arrays");
        any = true;
    }
    if (!any) throw std::invalid_argument("no output given: pass at least one of A, B, E, G");
    py::gil_scoped_release nogil;
    src.computeAllFieldsGPU(p.ptr, (size_t)p.rows, out[0].ptr, out[1].ptr, out[2].ptr, out[3].ptr, p.device);
}

// Host arrays in, a new numpy array out: q selects the quantity (0 A, 1 B, 2 E, 3 G).
template <class Source>
py::array compute(const Source &src, py::handle points, int q, ComputeMethod m, const char *what)
{
    if (m != GPU) cpuOnly(what);
    Buffer p = view(points, 3, false, "point_vectors");
    if (p.device) throw std::invalid_argument("device arrays: use compute_all_fields_into with device outputs");
    py::array_t<double> out = q == 3 ? py::array_t<double>({(py::ssize_t)p.rows, (py::ssize_t)3, (py::ssize_t)3})
                                     : py::array_t<double>({(py::ssize_t)p.rows, (py::ssize_t)3});
    double *o = out.mutable_data();
    {
        py::gil_scoped_release nogil;
        src.computeAllFieldsGPU(p.ptr, (size_t)p.rows, q == 0 ? o : nullptr, q == 1 ? o : nullptr,
                                q == 2 ? o : nullptr, q == 3 ? o : nullptr, false);
    }
    return out;
}

template <class Cls, class Source>
void add_methods(Cls &cls, const char *owner)
{
    static const std::string names[4] = {std::string(owner) + ".compute_all_A_potential_vectors",
                                         std::string(owner) + ".compute_all_B_field_vectors",
                                         std::string(owner) + ".compute_all_E_field_vectors",
                                         std::string(owner) + ".compute_all_B_gradient_matrices"};
    const char *py_names[4] = {"compute_all_A_potential_vectors", "compute_all_B_field_vectors",
                               "compute_all_E_field_vectors", "compute_all_B_gradient_matrices"};
    for (int q = 0; q < 4; ++q)
        cls.def(py_names[q],
                [q](const Source &self, py::array points, ComputeMethod m) {
                    return compute(self, points, q, m, names[q].c_str());
                },
                py::arg("point_vectors").noconvert(), py::arg("compute_method") = CPU_ST);
    cls.def("compute_all_fields_into",
            [](const Source &self, py::object points, py::object A, py::object B, py::object E, py::object G) {
                into(self, points, A, B, E, G);
            },
            py::arg("points"), py::arg("A") = py::none(), py::arg("B") = py::none(), py::arg("E") = py::none(),
            py::arg("G") = py::none(),
            "Fields at the GPU precision into caller buffers (numpy arrays, or DLPack arrays such as torch tensors on\n"
            "the CPU or the GPU; all host or all device), every requested quantity in one launch. points: (n, 3);\n"
            "A, B, E: (n, 3); G: (n, 9) or (n, 3, 3), the gradient in the reference's convention.");
}

}  // namespace zc

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
        .def_static("set_gpu_far_field", &Coil::setGpuFarField, py::arg("mode"))
        .def_static("get_gpu_far_field", &Coil::getGpuFarField)
        .def("__repr__", &Coil::operator std::string);
    zc::add_methods<decltype(coil), Coil>(coil, "Coil");

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
    py::class_<CoilGroup> group(g, "CoilGroup");
    group.def(py::init<std::vector<std::shared_ptr<Coil>>, PrecisionFactor, int>(),
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
    zc::add_methods<decltype(group), CoilGroup>(group, "CoilGroup");
}

PYBIND11_MODULE(_c_coil, m)
{
    m.doc() = "C-Coil (MI355X build): circular-coil magnetostatics on gfx950 through the ccoil_hip C ABI";
    init_tensor(m);
    init_coil(m);
    init_coil_group(m);
    m.attr("__version__") = ccoil_version();
}
