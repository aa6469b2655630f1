// vec3 tensor types of the C-Coil C++ API (interface of src/Tensor/Tensor.h:27-253, kept so user code
// compiles unchanged). Vector3Array / Matrix3Array store AoS doubles, which is also the point/output
// layout of the C ABI, so arrays are handed to the GPU path without repacking.
#pragma once

#include <string>
#include <vector>

namespace vec3 {

class Triplet {
  public:
    double first, second, third;
    Triplet();
    explicit Triplet(double first, double second, double third);
    explicit operator std::string() const;
};

class Vector3 {
  public:
    double x, y, z;
    Vector3();
    explicit Vector3(double x, double y, double z);

    Vector3 operator+(const Vector3 &o) const;
    Vector3 operator+=(const Vector3 &o);
    Vector3 operator-(const Vector3 &o) const;
    Vector3 operator-=(const Vector3 &o);
    Vector3 operator*(double m) const;
    Vector3 operator*=(double m);

    [[nodiscard]] double abs() const;
    static double scalarProduct(Vector3 a, Vector3 b);
    static Vector3 crossProduct(Vector3 a, Vector3 b);
    static Vector3 getFromCylindricalCoords(double z, double r, double phi);
    static Vector3 getFromSphericalCoords(double r, double theta, double phi);
    [[nodiscard]] Triplet getAsCylindricalCoords() const;
    [[nodiscard]] Triplet getAsSphericalCoords() const;
    explicit operator std::string() const;
};

class Matrix3 {
  public:
    double xx, xy, xz, yx, yy, yz, zx, zy, zz;
    Matrix3();
    explicit Matrix3(double xx, double xy, double xz, double yx, double yy, double yz, double zx, double zy,
                     double zz);
    [[nodiscard]] double det() const;
    Matrix3 operator+(const Matrix3 &m) const;
    Matrix3 operator+=(const Matrix3 &m);
    void operator*=(double m);
    Matrix3 operator*(double m) const;
    Matrix3 operator*(const Matrix3 &m) const;
    Vector3 operator*(const Vector3 &v) const;
    explicit operator std::string() const;
};

class Vector3Array {
    std::vector<Vector3> items_;

  public:
    Vector3Array();
    explicit Vector3Array(size_t initSize);
    explicit Vector3Array(const std::vector<Vector3> &items);

    void append(const Vector3 &v);
    void append(double x, double y, double z);
    void reserve(size_t n);
    void resize(size_t n);
    void clear();
    [[nodiscard]] size_t size() const;
    std::vector<Vector3> &getItems();
    [[nodiscard]] const std::vector<Vector3> &getItems() const;

    [[nodiscard]] std::vector<double> x() const;
    [[nodiscard]] std::vector<double> y() const;
    [[nodiscard]] std::vector<double> z() const;
    [[nodiscard]] std::vector<double> abs() const;

    Vector3 &operator[](size_t i);
    const Vector3 &operator[](size_t i) const;
    Vector3Array &operator+=(const Vector3 &v);
    explicit operator std::string() const;

    // Contiguous AoS view (3 doubles per element).
    [[nodiscard]] const double *data() const { return reinterpret_cast<const double *>(items_.data()); }
    double *data() { return reinterpret_cast<double *>(items_.data()); }
};

class Matrix3Array {
    std::vector<Matrix3> items_;

  public:
    Matrix3Array();
    explicit Matrix3Array(size_t initSize);
    explicit Matrix3Array(const std::vector<Matrix3> &items);

    void append(const Matrix3 &m);
    void append(double xx, double xy, double xz, double yx, double yy, double yz, double zx, double zy, double zz);
    void reserve(size_t n);
    void resize(size_t n);
    void clear();
    [[nodiscard]] size_t size() const;
    std::vector<Matrix3> &getItems();
    [[nodiscard]] const std::vector<Matrix3> &getItems() const;

    [[nodiscard]] std::vector<double> xx() const;
    [[nodiscard]] std::vector<double> xy() const;
    [[nodiscard]] std::vector<double> xz() const;
    [[nodiscard]] std::vector<double> yx() const;
    [[nodiscard]] std::vector<double> yy() const;
    [[nodiscard]] std::vector<double> yz() const;
    [[nodiscard]] std::vector<double> zx() const;
    [[nodiscard]] std::vector<double> zy() const;
    [[nodiscard]] std::vector<double> zz() const;
    [[nodiscard]] std::vector<double> det() const;

    Matrix3 &operator[](size_t i);
    const Matrix3 &operator[](size_t i) const;
    Matrix3Array &operator+=(const Matrix3 &m);
    explicit operator std::string() const;

    [[nodiscard]] const double *data() const { return reinterpret_cast<const double *>(items_.data()); }
    double *data() { return reinterpret_cast<double *>(items_.data()); }
};

static_assert(sizeof(Vector3) == 3 * sizeof(double), "Vector3 must be three packed doubles");
static_assert(sizeof(Matrix3) == 9 * sizeof(double), "Matrix3 must be nine packed doubles");

}  // namespace vec3
