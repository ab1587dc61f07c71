// Host-side model of a 2-D mocc problem: materials, pin meshes, the core
// mesh, the angular quadrature and the ray tracer. Everything here runs once
// at set-up on the CPU; its output is the structure-of-arrays description in
// include/mocgpu.h that the device sweeper and the oracle both consume.
//
// The geometric arithmetic (intersections, sorting, tie-breaking) follows the
// reference operation for operation, because its golden ray tests compare
// segment lengths at 1e-15 (src/sweepers/moc/tests/test_Ray.cpp:281-305).
// One step is added: CoreMesh::snap_corners merges the two crossings of a ray
// that grazes a pin corner, where the reference's fuzzy point equality and
// corner classification disagree and count one coarse surface twice. Rays
// that pass no closer than 1e-9 cm to a corner are traced exactly as before.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "xml.hpp"

namespace mocc_amd {

using VecF = std::vector<double>;
using VecI = std::vector<int>;

constexpr double kPI = 3.1415926535897932;
constexpr double kTWOPI = 2.0 * kPI;
constexpr double kHPI = 0.5 * kPI;
constexpr double kFPI = 4.0 * kPI;
constexpr double kREAL_FUZZ = 200 * 2.220446049250313080847e-16;

enum Surface { EAST = 0, NORTH = 1, WEST = 2, SOUTH = 3, TOP = 4, BOTTOM = 5, INTERNAL = 6, INVALID = 7 };
enum class Normal : int { X = 0, Y = 1, Z = 2 };
enum class Boundary : int { VACUUM = 0, REFLECT = 1, PARALLEL = 2, PERIODIC = 3, PRESCRIBED = 4, INVALID = 5 };

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

// --- floating-point comparisons (util/fp_utils.hpp:39-126) -------------------
bool fp_equiv_ulp(double a, double b);
bool fp_equiv_abs(double a, double b);
bool fp_equiv(double a, double b);
inline bool fuzzy_lt(double l, double r) { return (l - r) / std::abs(r) < -kREAL_FUZZ; }

// --- plane geometry (core/geometry/*) ------------------------------------------
struct Point2 {
    double x = 0.0, y = 0.0;
    bool ok = false;
    Point2() = default;
    Point2(double x_, double y_) : x(x_), y(y_), ok(true) {}
    double distance(const Point2 &p) const { return std::sqrt((x - p.x) * (x - p.x) + (y - p.y) * (y - p.y)); }
    bool operator<(const Point2 &o) const { return y < o.y; }
    bool operator==(const Point2 &o) const { return fp_equiv(x, o.x) && fp_equiv(y, o.y); }
    Point2 operator-(const Point2 &o) const { return Point2(x - o.x, y - o.y); }
    double alpha() const { return (y > 0.0) ? std::atan2(y, x) : std::atan2(y, x) + kTWOPI; }
};
inline Point2 Midpoint(const Point2 &a, const Point2 &b) { return Point2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y)); }

struct Line {
    Point2 p1, p2;
    Line(Point2 a, Point2 b) : p1(a), p2(b) {}
};
struct Circle {
    Point2 c;
    double r;
};
int intersect_line_circle(const Line &l, const Circle &c, Point2 &p1, Point2 &p2);
int intersect_lines(const Line &l1, const Line &l2, Point2 &p);

struct Direction {
    double ox, oy, oz, alpha, theta, rsintheta;
    Direction() = default;
    static Direction from_angles(double alpha, double theta);
    static Direction from_cosines(double ox, double oy, double oz);
    Direction to_octant(int octant) const;
};

struct Angle : Direction {
    double weight = 0.0;
    Angle() = default;
    Angle(const Direction &d, double w) : Direction(d), weight(w) {}
    static Angle make(double alpha, double theta, double w) { return Angle(Direction::from_angles(alpha, theta), w); }
    Angle to_octant(int octant) const { return Angle(Direction::to_octant(octant), weight); }
    void modify_alpha(double a) { Direction d = Direction::from_angles(a, theta); static_cast<Direction &>(*this) = d; }
};

// Box of the core, Box::intersect (core/geometry/box.cpp).
Point2 box_intersect(const Point2 &lo, const Point2 &hi, const Point2 &p, double alpha);

// --- angular quadrature (core/angular_quadrature.cpp, product_quadrature.hpp,
//     level_symmetric.hpp) -------------------------------------------------------
enum class QuadType { LS, CHEB_GAUSS, CHEB_YAMAMOTO, USER };

std::vector<std::pair<double, double>> gen_chebyshev(int n_azi);
std::vector<std::pair<double, double>> gen_yamamoto(int n_pol);
std::vector<std::pair<double, double>> gen_gauss(int n_pol);
std::vector<Angle> gen_product(const std::vector<std::pair<double, double>> &azi,
                               const std::vector<std::pair<double, double>> &pol);
std::vector<Angle> gen_level_symmetric(int order);

class AngularQuadrature {
public:
    AngularQuadrature() = default;
    explicit AngularQuadrature(const XmlNode &ang_quad);
    int ndir_oct() const { return ndir_oct_; }
    int ndir() const { return (int)angles_.size(); }
    int n_polar() const { return n_polar_; }
    int n_azimuthal() const { return n_azimuthal_; }
    QuadType type() const { return type_; }
    const Angle &operator[](int i) const { return angles_[i]; }
    int reflect(int iang, Normal n) const;
    int reverse(int iang) const { return (iang + ndir_oct_ * 2) % (ndir_oct_ * 4); }
    void modify_angle(int iang, const Angle &a);
    void update_weights();

private:
    void expand_octants();
    QuadType type_ = QuadType::LS;
    int ndir_oct_ = 0;
    int n_polar_ = -1, n_azimuthal_ = -1;
    std::vector<Angle> angles_;
};

// --- materials (core/material.cpp, core/material_lib.cpp) ------------------------
struct Material {
    std::string name;
    VecF xsab, xstr, xsnf, xsf, xsch;
    std::vector<VecF> scat;  // [to][from]
    bool fissile = false;
};

struct MaterialLib {
    int n_group = 0;
    VecF g_bounds;
    std::vector<Material> lib;
    std::map<std::string, int> by_name;
    std::map<int, int> id_to_lib;  // user material id -> lib index
    std::map<int, std::string> id_to_name;
    void load(const XmlNode &material_lib, const std::string &base_dir);
    // Re-read the library from another XSL file with the same material names
    // (cross-section branch cases); the deck's id -> name binding is kept.
    void reload(const std::string &path);
    const Material &by_id(int id) const;
    bool has(int id) const { return id_to_lib.count(id) > 0; }
};

// --- pin meshes (core/pin_mesh_cyl.cpp, core/pin_mesh_rect.cpp) --------------
class PinMesh {
public:
    enum Kind { CYL, RECT };
    static std::unique_ptr<PinMesh> from_xml(const XmlNode &n);
    int id = 0;
    Kind kind = RECT;
    double pitch_x = 0, pitch_y = 0;
    int n_reg = 0, n_xsreg = 0;
    VecF areas;
    // cyl
    VecF xs_radii, radii;
    VecI sub_rad;
    int sub_azi = 0;
    std::vector<Circle> circles;
    std::vector<Line> lines;
    // rect
    int nxr = 0, nyr = 0;
    VecF hx, hy;

    int n_fsrs(int xsreg) const;
    int find_reg(const Point2 &p) const;
    // Trace p1 -> p2 (pin-local); appends lengths and region indices.
    int trace(const Point2 &p1, const Point2 &p2, int first_reg, VecF &s, VecI &reg) const;
};

struct Pin {
    int id = 0;
    const PinMesh *mesh = nullptr;
    VecI mat_ids;
    bool is_fuel = false;
};

struct Lattice {
    int id = 0;
    int nx = 0, ny = 0;
    std::vector<const Pin *> pins;  // row-major from the lower-left
    VecF hx_vec, hy_vec, x_vec, y_vec;
    double hx = 0, hy = 0;
    int n_reg = 0;
    VecI first_reg_pin;
    const Pin &at(int ix, int iy) const { return *pins[iy * nx + ix]; }
};

// One 2-D plane of the core: a CoreMesh restricted to what the MoC path needs.
class CoreMesh {
public:
    explicit CoreMesh(const XmlNode &doc, const std::string &base_dir);
    int nx = 0, ny = 0;               // pins
    int nlat_x = 0, nlat_y = 0;       // lattices
    double hx_core = 0, hy_core = 0, hz = 0;
    VecF x_vec, y_vec;                // global pin boundaries
    VecF lat_x, lat_y;                // lattice boundaries (Plane::hx_)
    std::array<Boundary, 6> bc{};
    MaterialLib mat_lib;
    std::map<int, std::unique_ptr<PinMesh>> pin_meshes;
    std::map<int, std::unique_ptr<Pin>> pins;
    std::map<int, std::unique_ptr<Lattice>> lattices;
    std::vector<const Lattice *> plane_lattices;  // [nlat_y*nlat_x]
    VecI first_reg_lattice;
    std::vector<const Pin *> core_pins;           // FSR order
    VecI pin_first_reg;                           // FSR offset of core_pins[i]
    int n_reg = 0;
    int n_fuel = 0;
    std::vector<Line> lines;                      // internal pin boundaries

    // Coarse-mesh helpers (core/mesh.cpp, core/mesh.hpp) for one plane.
    int n_surf_plane() const { return (nx + 1) * ny + (ny + 1) * nx + nx * ny; }
    int n_surf() const { return (nx + 1) * ny + (ny + 1) * nx + 2 * nx * ny; }
    int coarse_cell(int ix, int iy) const { return iy * nx + ix; }
    int coarse_surf(int cell, Surface s) const;
    int coarse_neighbor(int cell, Surface s) const;
    size_t coarse_norm_point(const Point2 &p, int octant, Surface (&s)[2]) const;
    int coarse_boundary_cell(const Point2 &p, int octant) const;
    int pin_position_x(int ipin) const;
    int pin_position_y(int ipin) const;
    // CoreMesh::get_pinmesh: locate the pin holding p, return its mesh, set
    // p to the pin centre and add the pin's first region to first_reg.
    const PinMesh *get_pinmesh(Point2 &p, int &first_reg) const;
    // Trace the core lines (Mesh::trace, core/mesh.cpp:321-344), then merge
    // the two crossings of a ray that grazes a pin corner.
    void trace(std::vector<Point2> &ps) const;
    void snap_corners(std::vector<Point2> &ps) const;
    // distance (cm) below which a crossing is taken to pass through a corner
    static constexpr double kCornerSnap = 1e-9;
    VecF areas() const;
};

// --- rays (sweepers/moc/ray.cpp, ray_data.cpp) -----------------------------------
struct RayCoarseData {
    int fw, bw, nseg_fw, nseg_bw;
};

struct Ray {
    Point2 p1, p2;
    int bc[2] = {0, 0};
    VecF seg_len;
    VecI seg_index;
    int cm_cell_fw = 0, cm_cell_bw = 0, cm_surf_fw = 0, cm_surf_bw = 0;
    std::vector<RayCoarseData> cm_data;
    Ray(Point2 a, Point2 b, int bc0, int bc1, const CoreMesh &mesh);
    int nseg() const { return (int)seg_len.size(); }
};

enum class VolumeCorrection { FLAT, ANGLE, NONE };

class RayData {
public:
    RayData(const XmlNode &rays, const AngularQuadrature &quad, const CoreMesh &mesh);
    AngularQuadrature ang_quad;       // modularized
    VecI Nx, Ny;                      // per octant-1/2 angle
    VecF spacing;
    // rays_[a] for azimuth a over octants 1-2 (shared by the polar angles)
    std::vector<std::vector<Ray>> rays;
    int max_seg = 0;
    VolumeCorrection correction = VolumeCorrection::FLAT;

private:
    std::pair<int, int> modularize_angle(const Angle &ang, double hx, double hy, double spacing) const;
    void correct_volume(const CoreMesh &mesh);
};

std::pair<int, int> rational_approximation(double target, double tol, int max_d);

// Host threads for the tracer and the array assembly: MOCC_HOST_THREADS, else
// OMP_NUM_THREADS, else the hardware's count, within [1, 64].
int host_threads();

// f(i) for every i in [0, n), spread over host_threads() threads; each i runs
// exactly once and writes only its own outputs, so the result does not depend
// on the thread count. The first exception thrown is rethrown here after all
// threads have stopped.
void parallel_for(int n, const std::function<void(int)> &f);

}  // namespace mocc_amd
