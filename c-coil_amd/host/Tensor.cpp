#include "Tensor.h"

#include <cmath>
#include <sstream>

namespace vec3 {

namespace {
template <typename T> std::string join(const std::vector<T> &v, const char *name)
{
    std::ostringstream o;
    o << name << "([";
    for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << std::string(v[i]);
    o << "])";
    return o.str();
}
}  // namespace

Triplet::Triplet() : first(0), second(0), third(0) {}
Triplet::Triplet(double f, double s, double t) : first(f), second(s), third(t) {}
Triplet::operator std::string() const
{
    std::ostringstream o;
    o << "Triplet(first=" << first << ", second=" << second << ", third=" << third << ")";
    return o.str();
}

Vector3::Vector3() : x(0), y(0), z(0) {}
Vector3::Vector3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
Vector3 Vector3::operator+(const Vector3 &o) const { return Vector3(x + o.x, y + o.y, z + o.z); }
Vector3 Vector3::operator+=(const Vector3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
Vector3 Vector3::operator-(const Vector3 &o) const { return Vector3(x - o.x, y - o.y, z - o.z); }
Vector3 Vector3::operator-=(const Vector3 &o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
Vector3 Vector3::operator*(double m) const { return Vector3(x * m, y * m, z * m); }
Vector3 Vector3::operator*=(double m) { x *= m; y *= m; z *= m; return *this; }
double Vector3::abs() const { return std::sqrt(x * x + y * y + z * z); }
double Vector3::scalarProduct(Vector3 a, Vector3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vector3 Vector3::crossProduct(Vector3 a, Vector3 b)
{
    return Vector3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
Vector3 Vector3::getFromCylindricalCoords(double z, double r, double phi)
{
    return Vector3(r * std::cos(phi), r * std::sin(phi), z);
}
Vector3 Vector3::getFromSphericalCoords(double r, double theta, double phi)
{
    const double rho = r * std::sin(theta);
    return Vector3(rho * std::cos(phi), rho * std::sin(phi), r * std::cos(theta));
}
Triplet Vector3::getAsCylindricalCoords() const { return Triplet(z, std::hypot(x, y), std::atan2(y, x)); }
Triplet Vector3::getAsSphericalCoords() const
{
    return Triplet(std::sqrt(x * x + y * y + z * z), std::atan2(std::hypot(x, y), z), std::atan2(y, x));
}
Vector3::operator std::string() const
{
    std::ostringstream o;
    o << "Vector3(x=" << x << ", y=" << y << ", z=" << z << ")";
    return o.str();
}

Matrix3::Matrix3() : xx(0), xy(0), xz(0), yx(0), yy(0), yz(0), zx(0), zy(0), zz(0) {}
Matrix3::Matrix3(double a, double b, double c, double d, double e, double f, double g, double h, double i)
    : xx(a), xy(b), xz(c), yx(d), yy(e), yz(f), zx(g), zy(h), zz(i) {}
double Matrix3::det() const
{
    return xx * (yy * zz - yz * zy) - xy * (yx * zz - yz * zx) + xz * (yx * zy - yy * zx);
}
Matrix3 Matrix3::operator+(const Matrix3 &m) const
{
    return Matrix3(xx + m.xx, xy + m.xy, xz + m.xz, yx + m.yx, yy + m.yy, yz + m.yz, zx + m.zx, zy + m.zy, zz + m.zz);
}
Matrix3 Matrix3::operator+=(const Matrix3 &m) { *this = *this + m; return *this; }
void Matrix3::operator*=(double m) { *this = *this * m; }
Matrix3 Matrix3::operator*(double m) const
{
    return Matrix3(xx * m, xy * m, xz * m, yx * m, yy * m, yz * m, zx * m, zy * m, zz * m);
}
Matrix3 Matrix3::operator*(const Matrix3 &m) const
{
    const double a[9] = {xx, xy, xz, yx, yy, yz, zx, zy, zz};
    const double b[9] = {m.xx, m.xy, m.xz, m.yx, m.yy, m.yz, m.zx, m.zy, m.zz};
    double c[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    return Matrix3(c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8]);
}
Vector3 Matrix3::operator*(const Vector3 &v) const
{
    return Vector3(xx * v.x + xy * v.y + xz * v.z, yx * v.x + yy * v.y + yz * v.z, zx * v.x + zy * v.y + zz * v.z);
}
Matrix3::operator std::string() const
{
    std::ostringstream o;
    o << "Matrix3([[" << xx << ", " << xy << ", " << xz << "], [" << yx << ", " << yy << ", " << yz << "], [" << zx
      << ", " << zy << ", " << zz << "]])";
    return o.str();
}

Vector3Array::Vector3Array() = default;
Vector3Array::Vector3Array(size_t n) : items_(n) {}
Vector3Array::Vector3Array(const std::vector<Vector3> &items) : items_(items) {}
void Vector3Array::append(const Vector3 &v) { items_.push_back(v); }
void Vector3Array::append(double x, double y, double z) { items_.emplace_back(x, y, z); }
void Vector3Array::reserve(size_t n) { items_.reserve(n); }
void Vector3Array::resize(size_t n) { items_.resize(n); }
void Vector3Array::clear() { items_.clear(); }
size_t Vector3Array::size() const { return items_.size(); }
std::vector<Vector3> &Vector3Array::getItems() { return items_; }
const std::vector<Vector3> &Vector3Array::getItems() const { return items_; }
std::vector<double> Vector3Array::x() const { std::vector<double> o; o.reserve(size()); for (auto &v : items_) o.push_back(v.x); return o; }
std::vector<double> Vector3Array::y() const { std::vector<double> o; o.reserve(size()); for (auto &v : items_) o.push_back(v.y); return o; }
std::vector<double> Vector3Array::z() const { std::vector<double> o; o.reserve(size()); for (auto &v : items_) o.push_back(v.z); return o; }
std::vector<double> Vector3Array::abs() const { std::vector<double> o; o.reserve(size()); for (auto &v : items_) o.push_back(v.abs()); return o; }
Vector3 &Vector3Array::operator[](size_t i) { return items_[i]; }
const Vector3 &Vector3Array::operator[](size_t i) const { return items_[i]; }
Vector3Array &Vector3Array::operator+=(const Vector3 &v) { items_.push_back(v); return *this; }
Vector3Array::operator std::string() const { return join(items_, "Vector3Array"); }

Matrix3Array::Matrix3Array() = default;
Matrix3Array::Matrix3Array(size_t n) : items_(n) {}
Matrix3Array::Matrix3Array(const std::vector<Matrix3> &items) : items_(items) {}
void Matrix3Array::append(const Matrix3 &m) { items_.push_back(m); }
void Matrix3Array::append(double a, double b, double c, double d, double e, double f, double g, double h, double i)
{
    items_.emplace_back(a, b, c, d, e, f, g, h, i);
}
void Matrix3Array::reserve(size_t n) { items_.reserve(n); }
void Matrix3Array::resize(size_t n) { items_.resize(n); }
void Matrix3Array::clear() { items_.clear(); }
size_t Matrix3Array::size() const { return items_.size(); }
std::vector<Matrix3> &Matrix3Array::getItems() { return items_; }
const std::vector<Matrix3> &Matrix3Array::getItems() const { return items_; }
#define CCOIL_M3_COMPONENT(name)                                                       \
    std::vector<double> Matrix3Array::name() const                                     \
    {                                                                                  \
        std::vector<double> o;                                                         \
        o.reserve(size());                                                             \
        for (auto &m : items_) o.push_back(m.name);                                    \
        return o;                                                                      \
    }
CCOIL_M3_COMPONENT(xx)
CCOIL_M3_COMPONENT(xy)
CCOIL_M3_COMPONENT(xz)
CCOIL_M3_COMPONENT(yx)
CCOIL_M3_COMPONENT(yy)
CCOIL_M3_COMPONENT(yz)
CCOIL_M3_COMPONENT(zx)
CCOIL_M3_COMPONENT(zy)
CCOIL_M3_COMPONENT(zz)
#undef CCOIL_M3_COMPONENT
std::vector<double> Matrix3Array::det() const { std::vector<double> o; o.reserve(size()); for (auto &m : items_) o.push_back(m.det()); return o; }
Matrix3 &Matrix3Array::operator[](size_t i) { return items_[i]; }
const Matrix3 &Matrix3Array::operator[](size_t i) const { return items_[i]; }
Matrix3Array &Matrix3Array::operator+=(const Matrix3 &m) { items_.push_back(m); return *this; }
Matrix3Array::operator std::string() const { return join(items_, "Matrix3Array"); }

}  // namespace vec3
