#include "model.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <fstream>
#include <limits>
#include <numeric>
#include <regex>
#include <sstream>

namespace mocc_amd {

// ---------------------------------------------------------------------------
// floating-point comparisons, util/fp_utils.hpp:39-126
// ---------------------------------------------------------------------------
bool fp_equiv_ulp(double a, double b)
{
    int64_t i1, i2;
    std::memcpy(&i1, &a, 8);
    std::memcpy(&i2, &b, 8);
    if (i1 < 0) i1 = 2147483648LL - i1;
    if (i2 < 0) i2 = 2147483648LL - i2;
    int64_t d = i1 - i2;
    if (d < 0) d = -d;
    return d < 200;
}
bool fp_equiv_abs(double a, double b) { return std::abs(a - b) < kREAL_FUZZ; }
bool fp_equiv(double a, double b)
{
    if (fp_equiv_abs(a, b)) return true;
    if ((a < 0.0) != (b < 0.0)) return false;
    return fp_equiv_ulp(a, b);
}

namespace {
template <typename T> std::vector<T> explode(const std::string &data)
{
    std::vector<T> out;
    std::istringstream in(data);
    std::string tok;
    while (in >> tok) {
        std::istringstream t(tok);
        T v;
        t >> v;
        if (t.fail()) throw Error("Malformed data: " + data);
        out.push_back(v);
    }
    return out;
}

// util/string_utils.cpp explode_braces: "{1 2} 3" -> {{1,2},{3}}
std::vector<VecI> explode_braces(const std::string &data)
{
    std::vector<VecI> out;
    size_t pos = 0;
    while (pos < data.size()) {
        size_t stt = data.find('{', pos);
        std::string naked = data.substr(pos, stt == std::string::npos ? std::string::npos : stt - pos);
        for (int v : explode<int>(naked)) out.push_back(VecI(1, v));
        if (stt == std::string::npos) break;
        size_t stp = data.find('}', stt);
        if (stp == std::string::npos) throw Error("Brace mismatch");
        VecI b = explode<int>(data.substr(stt + 1, stp - stt - 1));
        if (b.empty()) throw Error("Empty brackets");
        out.push_back(b);
        pos = stp + 1;
    }
    return out;
}
}  // namespace

// ---------------------------------------------------------------------------
// plane geometry, core/geometry/geom.hpp
// ---------------------------------------------------------------------------
int intersect_line_circle(const Line &l, const Circle &circ, Point2 &p1, Point2 &p2)
{
    int ret = 0;
    double u1 = l.p2.x - l.p1.x;
    double u2 = l.p2.y - l.p1.y;
    double w1 = l.p1.x - circ.c.x;
    double w2 = l.p1.y - circ.c.y;
    double b = w1 * u1 + w2 * u2;
    double c = w1 * w1 + w2 * w2 - circ.r * circ.r;
    if ((c > 0.0) & (b > 0.0)) {
        p1.ok = p2.ok = false;
        return 0;
    }
    double a = u1 * u1 + u2 * u2;
    double disc = b * b - a * c;
    if (disc < 0.0) {
        p1.ok = p2.ok = false;
        return 0;
    }
    // fp_equiv_rel(disc, 0.0) is |disc|/|disc| < fuzz: never true.
    double ra = 1.0 / a;
    disc = std::sqrt(disc);
    double t1 = (-b - disc) * ra;
    double t2 = (-b + disc) * ra;
    if ((0.0 < t1) & (t1 < 1.0)) {
        p1 = l.p1;
        p1.x += u1 * t1;
        p1.y += u2 * t1;
        p1.ok = true;
        ret++;
    }
    if ((0.0 < t2) & (t2 < 1.0)) {
        if (ret == 0) {
            p1 = l.p1;
            p1.x += u1 * t2;
            p1.y += u2 * t2;
            p1.ok = true;
        } else {
            p2 = l.p1;
            p2.x += u1 * t2;
            p2.y += u2 * t2;
            p2.ok = true;
        }
        ret++;
    }
    return ret;
}

int intersect_lines(const Line &l1, const Line &l2, Point2 &p)
{
    double u1 = l1.p2.x - l1.p1.x;
    double u2 = l1.p2.y - l1.p1.y;
    double v1 = l2.p2.x - l2.p1.x;
    double v2 = l2.p2.y - l2.p1.y;
    double w1 = l1.p1.x - l2.p1.x;
    double w2 = l1.p1.y - l2.p1.y;
    double d = u1 * v2 - u2 * v1;
    double s = v1 * w2 - v2 * w1;
    double t = u1 * w2 - u2 * w1;
    s = s / d;
    t = t / d;
    if ((0.0 <= s) & (s <= 1.0) & (0.0 <= t) & (t <= 1.0)) {
        p = l1.p1;
        p.x += s * u1;
        p.y += s * u2;
        p.ok = true;
        return 1;
    }
    p.ok = false;
    return 0;
}

// core/geometry/direction.hpp:57-83 (long double trigonometry as there)
Direction Direction::from_angles(double alpha, double theta)
{
    Direction d;
    d.alpha = alpha;
    d.theta = theta;
    long double lt = theta, la = alpha;
    d.ox = (double)(std::sin(lt) * std::cos(la));
    d.oy = (double)(std::sin(lt) * std::sin(la));
    d.oz = (double)std::cos(lt);
    d.rsintheta = (double)(1.0 / std::sin(lt));
    return d;
}

Direction Direction::from_cosines(double ox, double oy, double oz)
{
    Direction d;
    d.ox = ox;
    d.oy = oy;
    d.oz = oz;
    long double ox_big = ox;
    long double oz_big = oz;
    long double theta_big = std::acos(oz_big);
    d.theta = (double)theta_big;
    d.alpha = (double)std::acos(ox_big / std::sin(theta_big));
    if (oy < 0.0) d.alpha = kTWOPI - d.alpha;
    d.rsintheta = (double)(1.0L / std::sin(theta_big));
    return d;
}

Direction Direction::to_octant(int octant) const
{
    double ax = std::fabs(ox), ay = std::fabs(oy), az = std::fabs(oz);
    switch (octant) {
    case 1: return from_cosines(ax, ay, az);
    case 2: return from_cosines(-ax, ay, az);
    case 3: return from_cosines(-ax, -ay, az);
    case 4: return from_cosines(ax, -ay, az);
    case 5: return from_cosines(ax, ay, -az);
    case 6: return from_cosines(-ax, ay, -az);
    case 7: return from_cosines(-ax, -ay, -az);
    case 8: return from_cosines(ax, -ay, -az);
    }
    throw Error("bad octant");
}

// core/geometry/box.cpp Box::intersect
Point2 box_intersect(const Point2 &lo, const Point2 &hi, const Point2 &p, double alpha)
{
    double ox = std::cos(alpha);
    double oy = std::sin(alpha);
    double d_min = 1.0e12;
    Point2 out;
    double x, y, d;
    x = lo.x;
    d = (lo.x - p.x) / ox;
    y = p.y + oy * d;
    if ((std::fabs(d) > kREAL_FUZZ) && (d < d_min) && (y > lo.y) && (y < hi.y)) {
        d_min = std::fabs(d);
        out = Point2(x, y);
    }
    x = hi.x;
    d = (hi.x - p.x) / ox;
    y = p.y + oy * d;
    if ((std::fabs(d) > kREAL_FUZZ) && (d < d_min) && (y > lo.y) && (y < hi.y)) {
        d_min = std::fabs(d);
        out = Point2(x, y);
    }
    y = lo.y;
    d = (lo.y - p.y) / oy;
    x = p.x + ox * d;
    if ((std::fabs(d) > kREAL_FUZZ) && (d < d_min) && (x > lo.x) && (x < hi.x)) {
        d_min = std::fabs(d);
        out = Point2(x, y);
    }
    y = hi.y;
    d = (hi.y - p.y) / oy;
    x = p.x + ox * d;
    if ((std::fabs(d) > kREAL_FUZZ) && (d < d_min) && (x > lo.x) && (x < hi.x)) {
        d_min = std::fabs(d);
        out = Point2(x, y);
    }
    return out;
}

// ---------------------------------------------------------------------------
// angular quadrature
// ---------------------------------------------------------------------------
// product_quadrature.hpp GenChebyshev: equally spaced azimuths in (0, pi/2)
std::vector<std::pair<double, double>> gen_chebyshev(int n)
{
    std::vector<std::pair<double, double>> out;
    double w = 1.0 / n;
    double del = 0.5 * kPI / (2 * n);
    for (int i = 0; i < n; i++) out.emplace_back(del * (2 * i + 1), w);
    return out;
}

// product_quadrature.hpp GenYamamoto: the published 3-angle Yamamoto set
// (polar angle, weight); only n_polar = 3 exists in the reference.
std::vector<std::pair<double, double>> gen_yamamoto(int n)
{
    if (n != 3) throw Error("Only support Yamamoto quadrature when npol=3");
    return {{0.167429147795000, 4.623300000000000E-002},
            {0.567715121084000, 0.283619000000000},
            {1.20253314678900, 0.670148000000000}};
}

// product_quadrature.hpp GenGauss: Gauss-Legendre on [-1, 1] with 2n nodes by
// Newton iteration from the Chebyshev-like initial guess; the n nodes in
// cos(theta) > 0 give the polar angles, weights normalized to sum to 1.
std::vector<std::pair<double, double>> gen_gauss(int n_polar)
{
    int N = n_polar * 2 - 1, N1 = n_polar * 2, N2 = n_polar * 2 + 1;
    VecF xu(N1), y(N1), w(N1), y0(N1, 2.0), Lp(N1);
    double delxu = 2.0 / N;
    for (int i = 0; i < N1; i++) {
        xu[i] = -1.0 + i * delxu;
        y[i] = std::cos((2 * i + 1) * kPI / (2 * N1)) + 0.27 / N1 * std::sin(kPI * xu[i] * N / N2);
    }
    std::vector<VecF> L(N1, VecF(N2));
    auto maxdiff = [&]() {
        double m = 0.0;
        for (int i = 0; i < N1; i++) m = std::max(m, std::abs(y[i] - y0[i]));
        return m;
    };
    while (maxdiff() > kREAL_FUZZ) {
        for (int m = 0; m < N1; m++) {
            L[m][0] = 1;
            L[m][1] = y[m];
        }
        for (int k = 1; k < N1; k++)
            for (int m = 0; m < N1; m++)
                L[m][k + 1] = ((2 * k + 1) * y[m] * L[m][k] - k * L[m][k - 1]) / (k + 1);
        for (int i = 0; i < N1; i++) Lp[i] = N2 * (L[i][N1 - 1] - y[i] * L[i][N2 - 1]) / (1 - y[i] * y[i]);
        y0 = y;
        for (int i = 0; i < N1; i++) y[i] = y0[i] - L[i][N2 - 1] / Lp[i];
    }
    for (int i = 0; i < N1; i++) w[i] = 1.0 / ((1 - y[i] * y[i]) * Lp[i] * Lp[i]) * N2 * N2 / (N1 * N1);
    for (int i = 0; i < N1; i++) y[i] = std::acos(y[i]);
    for (int i = 0; i < n_polar; i++) std::swap(y[i], y[N1 - 1 - i]);
    std::vector<std::pair<double, double>> out;
    for (int i = n_polar; i < N1; i++) out.emplace_back(y[i], w[i]);
    return out;
}

std::vector<Angle> gen_product(const std::vector<std::pair<double, double>> &azi,
                               const std::vector<std::pair<double, double>> &pol)
{
    std::vector<Angle> angles;
    double wsum = 0.0;
    for (const auto &a : azi) {
        for (const auto &p : pol) {
            angles.push_back(Angle::make(a.first, p.first, a.second * p.second));
            wsum += a.second * p.second;
        }
    }
    for (auto &a : angles) a.weight /= wsum;
    return angles;
}

// level_symmetric.hpp GenSn: the published level-symmetric base cosines and
// weight table (Lewis & Miller), octant 1, weights normalized to 1.
std::vector<Angle> gen_level_symmetric(int order)
{
    static const double mu_base[] = {0.577350269189626, 0.350021000000000, 0.266636000000000,
                                     0.218218218218218, 0.192450089729876, 0.174077655955702,
                                     0.161575000000000, 0.149071198499989};
    static const double w_unique[] = {
        1.0,       1.0 / 3.0, 0.1761262, 0.1572071, 0.1209876, 0.0907408, 0.0925925, 0.0893043,
        0.0725281, 0.0450455, 0.0539274, 0.0707734, 0.0558760, 0.0373436, 0.0502654, 0.0258553,
        0.0580031, 0.0488943, 0.0228095, 0.0393955, 0.0380920, 0.0258382, 0.0082759, 0.0489967,
        0.0413235, 0.0203158, 0.0265468, 0.0378883, 0.0135404, 0.0326129, 0.0103825};
    static const int w_offset[] = {0, 1, 2, 4, 7, 11, 16, 23};
    static const int w_map[] = {1, 1, 1, 1, 1, 2, 2, 1, 2, 1, 1, 2, 2, 2, 3, 2, 1, 2, 2, 1,
                                1, 2, 2, 3, 4, 3, 2, 4, 4, 2, 1, 2, 3, 2, 1, 1, 2, 2, 3, 5,
                                3, 4, 6, 6, 4, 3, 6, 7, 6, 3, 2, 5, 6, 6, 5, 2, 1, 2, 3, 4,
                                3, 2, 1, 1, 2, 2, 3, 5, 3, 4, 6, 6, 4, 4, 7, 8, 7, 4, 3, 6,
                                8, 8, 6, 3, 2, 5, 6, 7, 6, 5, 2, 1, 2, 3, 4, 4, 3, 2, 1};
    static const int w_map_offset[] = {0, 1, 4, 10, 20, 35, 63};
    if (order % 2 != 0) throw Error("Sn quadrature order must be even.");
    if (order > 16) throw Error("Max supported Sn quadrature order is 16.");
    const int n = order / 2;
    VecF mu{mu_base[n - 1]};
    if (order > 2) {
        const double delta_mu = 2.0 * (1.0 - 3.0 * (mu[0] * mu[0])) / (double)(order - 2);
        for (int i = 1; i < n; i++) mu.push_back(std::sqrt(mu[0] * mu[0] + i * delta_mu));
    }
    const double *weights = &w_unique[w_offset[n - 1]];
    const int *map = &w_map[w_map_offset[n - 1]];
    std::vector<Angle> angles;
    double wsum = 0.0;
    int k = 0;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j <= i; j++) {
            Angle a(Direction::from_cosines(mu[i - j], mu[j], mu[n - i - 1]), weights[map[k] - 1]);
            angles.push_back(a);
            wsum += a.weight;
            k++;
        }
    }
    for (auto &a : angles) a.weight /= wsum;
    return angles;
}

// angular_quadrature.cpp:47-113
AngularQuadrature::AngularQuadrature(const XmlNode &in)
{
    if (in.name != "ang_quad") throw Error("Input is not an <ang_quad/> tag");
    n_azimuthal_ = in.attr_int("n_azimuthal", -1);
    n_polar_ = in.attr_int("n_polar", -1);
    std::string t = in.attr_str("type");
    std::transform(t.begin(), t.end(), t.begin(), ::tolower);
    if (t == "ls" || t == "level-symmetric") {
        type_ = QuadType::LS;
        angles_ = gen_level_symmetric(in.attr_int("order", -1));
    } else if (t == "cg" || t == "chebyshev-gauss") {
        if (n_azimuthal_ < 1 || n_polar_ < 1) throw Error("Number of polar or azimuthal angles is invalid");
        type_ = QuadType::CHEB_GAUSS;
        angles_ = gen_product(gen_chebyshev(n_azimuthal_), gen_gauss(n_polar_));
    } else if (t == "cy" || t == "chebyshev-yamamoto") {
        if (n_azimuthal_ < 1 || n_polar_ < 1) throw Error("Number of polar or azimuthal angles is invalid");
        type_ = QuadType::CHEB_YAMAMOTO;
        angles_ = gen_product(gen_chebyshev(n_azimuthal_), gen_yamamoto(n_polar_));
    } else {
        throw Error("Invalid angular quadrature type specified: '" + t + "'");
    }
    ndir_oct_ = (int)angles_.size();
    expand_octants();
}

void AngularQuadrature::expand_octants()
{
    angles_.resize(ndir_oct_);
    for (int ioct = 2; ioct <= 8; ioct++)
        for (int i = 0; i < ndir_oct_; i++) angles_.push_back(angles_[i].to_octant(ioct));
}

int AngularQuadrature::reflect(int iang, Normal n) const
{
    static const int refl[3][8] = {{1, 0, 3, 2, 5, 4, 7, 6}, {3, 2, 1, 0, 7, 6, 5, 4}, {4, 5, 6, 7, 0, 1, 2, 3}};
    int ioct = iang / ndir_oct_;
    int new_oct = refl[(int)n][ioct];
    return iang + (new_oct - ioct) * ndir_oct_;
}

void AngularQuadrature::modify_angle(int iang, const Angle &a)
{
    angles_[iang] = a;
    for (int ioct = 1; ioct < 8; ioct++) angles_[iang + ioct * ndir_oct_] = a.to_octant(ioct + 1);
}

// angular_quadrature.cpp:246-352
void AngularQuadrature::update_weights()
{
    if (type_ != QuadType::CHEB_GAUSS && type_ != QuadType::CHEB_YAMAMOTO) return;
    auto polar = (type_ == QuadType::CHEB_GAUSS) ? gen_gauss(n_polar_) : gen_yamamoto(n_polar_);
    VecF azi;
    for (int i = 0; i < ndir_oct_; i++) azi.push_back(angles_[i].alpha);
    std::sort(azi.begin(), azi.end());
    azi.erase(std::unique(azi.begin(), azi.end(), fp_equiv_ulp), azi.end());
    if ((int)azi.size() != n_azimuthal_) throw Error("Wrong number of azimuthal angles!");
    VecF bounds{0.0};
    for (int i = 0; i < n_azimuthal_ - 1; i++) bounds.push_back((azi[i] + azi[i + 1]) / 2.0);
    bounds.push_back(kHPI);
    std::vector<std::pair<double, double>> pairs;
    for (int i = 0; i < n_azimuthal_; i++) {
        double hpi = kHPI;
        pairs.emplace_back(azi[i], (bounds[i + 1] - bounds[i]) / hpi);
    }
    angles_ = gen_product(pairs, polar);
    expand_octants();
}

// ---------------------------------------------------------------------------
// materials: core/material_lib.cpp:45-149, core/material.cpp:22-62
// ---------------------------------------------------------------------------
namespace {
void parse_xsl(MaterialLib &L, const std::string &path);
}

void MaterialLib::load(const XmlNode &in, const std::string &base_dir)
{
    std::string path = in.attr_str("path");
    if (path.empty()) throw Error("No material library specified.");
    if (path[0] != '/' && !base_dir.empty()) path = base_dir + "/" + path;
    parse_xsl(*this, path);
    for (const XmlNode *mn : in.children_named("material")) {
        int id = mn->attr_int("id", 0);
        std::string nm = mn->attr_str("name");
        auto it = by_name.find(nm);
        if (it == by_name.end()) throw Error("Material not found in library: " + nm);
        id_to_lib[id] = it->second;
        id_to_name[id] = nm;
    }
}

void MaterialLib::reload(const std::string &path)
{
    MaterialLib fresh;
    parse_xsl(fresh, path);
    for (const auto &kv : id_to_name) {
        auto it = fresh.by_name.find(kv.second);
        if (it == fresh.by_name.end()) throw Error("Material not found in library: " + kv.second);
        fresh.id_to_lib[kv.first] = it->second;
    }
    fresh.id_to_name = id_to_name;
    *this = std::move(fresh);
}

namespace {
void parse_xsl(MaterialLib &L, const std::string &path)
{
    int &n_group = L.n_group;
    VecF &g_bounds = L.g_bounds;
    auto &by_name = L.by_name;
    auto &lib = L.lib;
    std::ifstream f(path);
    if (!f) throw Error("Failed to open the cross-section library at: " + path);
    // FileScrubber: strip '!' comments and blank lines
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) {
        size_t c = line.find('!');
        if (c != std::string::npos) line.erase(c);
        size_t b = line.find_first_not_of(" \t\r\n");
        if (b == std::string::npos) continue;
        size_t e = line.find_last_not_of(" \t\r\n");
        lines.push_back(line.substr(b, e - b + 1));
    }
    size_t li = 0;
    auto next = [&]() -> std::string {
        if (li >= lines.size()) throw Error("Unexpected end of cross-section library");
        return lines[li++];
    };
    next();  // description
    int n_mat = 0;
    {
        std::istringstream b(next());
        b >> n_group >> n_mat;
        if (b.fail()) throw Error("Failed to read number of groups/materials!");
    }
    {
        std::istringstream b(next());
        for (int i = 0; i < n_group; i++) {
            double v;
            b >> v;
            if (b.fail()) throw Error("Trouble reading group bounds!");
            g_bounds.push_back(v);
        }
    }
    static const std::regex head("^\\s*XSMACRO\\s+([^\\s]+)\\s+([0-9]+)\\s*$");
    for (int im = 0; im < n_mat; im++) {
        std::smatch m;
        std::string h = next();
        if (!std::regex_match(h, m, head)) throw Error("Bad XSMACRO header: " + h);
        Material mat;
        mat.name = m[1].str();
        VecF ab(n_group), nf(n_group), fs(n_group), ch(n_group);
        for (int g = 0; g < n_group; g++) {
            VecF v = explode<double>(next());
            if (v.size() != 4) throw Error("Trouble reading XS data from library!");
            ab[g] = v[0];
            nf[g] = v[1];
            fs[g] = v[2];
            ch[g] = v[3];
        }
        mat.scat.assign(n_group, VecF(n_group, 0.0));
        for (int g = 0; g < n_group; g++) {
            VecF row = explode<double>(next());
            if ((int)row.size() != n_group) throw Error("Trouble reading scattering row");
            mat.scat[g] = row;
        }
        mat.xsab = ab;
        mat.xsnf = nf;
        mat.xsf = fs;
        mat.xsch = ch;
        mat.fissile = std::any_of(nf.begin(), nf.end(), [](double v) { return v > 0.0; });
        if (std::any_of(ch.begin(), ch.end(), [](double v) { return v > 0.0; })) {
            double s = 0.0;
            for (double c : mat.xsch) s += c;
            for (double &c : mat.xsch) c /= s;
        }
        // transport XS = absorption + out-scatter (ScatteringMatrix::out)
        VecF out(n_group, 0.0);
        for (int to = 0; to < n_group; to++)
            for (int from = 0; from < n_group; from++) out[from] += mat.scat[to][from];
        mat.xstr.resize(n_group);
        for (int g = 0; g < n_group; g++) mat.xstr[g] = ab[g] + out[g];
        by_name[mat.name] = im;
        lib.push_back(mat);
    }
}
}  // namespace

const Material &MaterialLib::by_id(int id) const
{
    auto it = id_to_lib.find(id);
    if (it == id_to_lib.end()) throw Error("Invalid material ID " + std::to_string(id));
    return lib[it->second];
}

// ---------------------------------------------------------------------------
// pin meshes
// ---------------------------------------------------------------------------
std::unique_ptr<PinMesh> PinMesh::from_xml(const XmlNode &n)
{
    auto pm = std::make_unique<PinMesh>();
    pm->id = n.attr_int("id", -1);
    if (pm->id < 0) throw Error("Failed to read pin ID.");
    pm->pitch_x = n.attr_double("pitch", -1.0);
    pm->pitch_y = pm->pitch_x;
    if (pm->pitch_x <= 0.0) throw Error("Failed to read valid pin pitch.");
    std::string type = n.attr_str("type");
    if (type == "cyl") {
        pm->kind = CYL;
        const XmlNode *r = n.child("radii");
        pm->xs_radii = explode<double>(r ? r->text : "");
        if (pm->xs_radii.empty()) throw Error("No radii for cylindrical pin mesh");
        for (size_t i = 0; i + 1 < pm->xs_radii.size(); i++)
            if (pm->xs_radii[i] > pm->xs_radii[i + 1]) throw Error("Pin radii do not appear to be ordered");
        if (pm->xs_radii.back() > pm->pitch_x * 0.5) throw Error("Largest radius is too big!");
        pm->n_xsreg = (int)pm->xs_radii.size() + 1;
        const XmlNode *az = n.child("sub_azi");
        VecI azv = explode<int>(az ? az->text : "");
        if (azv.empty()) throw Error("Improper input to azimuthal subdivisions!");
        pm->sub_azi = azv[0];
        if ((pm->sub_azi % 2 != 0) | (pm->sub_azi > 8)) throw Error("Only supporting even azimuthal subdivisions <=8.");
        const XmlNode *sr = n.child("sub_radii");
        pm->sub_rad = explode<int>(sr ? sr->text : "");
        if ((int)pm->sub_rad.size() != pm->n_xsreg - 1) throw Error("Wrong number of radial subdivisions specified.");
        // equal-volume rings within each XS ring (pin_mesh_cyl.cpp:114-128)
        double rxsi = 0.0, ri = 0.0;
        for (int ixs = 0; ixs < pm->n_xsreg - 1; ixs++) {
            double vn = (pm->xs_radii[ixs] * pm->xs_radii[ixs] - rxsi * rxsi) / pm->sub_rad[ixs];
            for (int ir = 0; ir < pm->sub_rad[ixs]; ir++) {
                double rr = std::sqrt(vn + ri * ri);
                pm->radii.push_back(rr);
                ri = rr;
            }
            rxsi = pm->xs_radii[ixs];
        }
        for (double rr : pm->radii) pm->circles.push_back(Circle{Point2(0.0, 0.0), rr});
        double hpx = 0.5 * pm->pitch_x, hpy = 0.5 * pm->pitch_y;
        double ang_sep = kTWOPI / pm->sub_azi;
        Point2 lo(-hpx, -hpy), hi(hpx, hpy), origin(0.0, 0.0);
        for (int ia = 0; ia < pm->sub_azi; ia++) {
            Point2 p = box_intersect(lo, hi, origin, ia * ang_sep);
            if (!p.ok) throw Error("Sector line does not reach the pin boundary");
            pm->lines.push_back(Line(origin, p));
        }
        pm->n_reg = ((int)pm->radii.size() + 1) * pm->sub_azi;
        double prev = 0.0;
        for (double rr : pm->radii) {
            double ai = kPI * (rr * rr - prev * prev) / pm->sub_azi;
            for (int ia = 0; ia < pm->sub_azi; ia++) pm->areas.push_back(ai);
            prev = rr;
        }
        double large = pm->radii.back();
        double v_out = (pm->pitch_x * pm->pitch_y - kPI * large * large) / pm->sub_azi;
        for (int ia = 0; ia < pm->sub_azi; ia++) pm->areas.push_back(v_out);
    } else if (type == "rect") {
        pm->kind = RECT;
        const XmlNode *sx = n.child("sub_x");
        const XmlNode *sy = n.child("sub_y");
        pm->nxr = sx ? std::atoi(sx->text.c_str()) : 0;
        pm->nyr = sy ? std::atoi(sy->text.c_str()) : 0;
        if (pm->nxr < 1 || pm->nyr < 1) throw Error("Failed to read valid number of divisions in rect pin mesh.");
        pm->n_xsreg = pm->n_reg = pm->nxr * pm->nyr;
        double dx = pm->pitch_x / pm->nxr, dy = pm->pitch_y / pm->nyr;
        double hpx = 0.5 * pm->pitch_x, hpy = 0.5 * pm->pitch_y;
        for (int i = 0; i <= pm->nxr; i++) pm->hx.push_back(i * dx - hpx);
        for (int i = 0; i <= pm->nyr; i++) pm->hy.push_back(i * dy - hpy);
        for (size_t ix = 1; ix + 1 < pm->hx.size(); ix++)
            pm->lines.push_back(Line(Point2(pm->hx[ix], -hpy), Point2(pm->hx[ix], hpy)));
        for (size_t iy = 1; iy + 1 < pm->hy.size(); iy++)
            pm->lines.push_back(Line(Point2(-hpx, pm->hy[iy]), Point2(hpx, pm->hy[iy])));
        pm->areas.assign(pm->n_reg, dx * dy);
    } else {
        throw Error("Unrecognized mesh type for mesh ID: " + std::to_string(pm->id));
    }
    return pm;
}

int PinMesh::n_fsrs(int xsreg) const
{
    if (kind == RECT) return 1;
    if (xsreg < (int)xs_radii.size()) return sub_rad[xsreg] * sub_azi;
    return sub_azi;
}

int PinMesh::find_reg(const Point2 &p) const
{
    if (std::fabs(p.x) > 0.5 * pitch_x || std::fabs(p.y) > 0.5 * pitch_y) return -1;
    if (kind == CYL) {
        double r = std::sqrt(p.x * p.x + p.y * p.y);
        size_t ir = 0;
        for (ir = 0; ir < radii.size(); ir++)
            if (r < radii[ir]) break;
        double azi = p.alpha();
        int ia = (int)(azi / (kTWOPI / sub_azi));
        return (int)ir * sub_azi + ia;
    }
    int ix = (int)(std::lower_bound(hx.begin(), hx.end(), p.x) - hx.begin()) - 1;
    int iy = (int)(std::lower_bound(hy.begin(), hy.end(), p.y) - hy.begin()) - 1;
    return nxr * iy + ix;
}

int PinMesh::trace(const Point2 &p1, const Point2 &p2, int first_reg, VecF &s, VecI &reg) const
{
    Line l(p1, p2);
    std::vector<Point2> ps{p1, p2};
    if (kind == CYL) {
        for (const auto &c : circles) {
            Point2 a, b;
            if (intersect_line_circle(l, c, a, b) == 2) {
                ps.push_back(a);
                ps.push_back(b);
            }
        }
    }
    for (const auto &li : lines) {
        Point2 p;
        if (intersect_lines(li, l, p) == 1) ps.push_back(p);
    }
    std::sort(ps.begin(), ps.end());
    ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
    for (size_t ip = 1; ip < ps.size(); ip++) {
        s.push_back(ps[ip].distance(ps[ip - 1]));
        int lr = find_reg(Midpoint(ps[ip], ps[ip - 1]));
        // A sliver along the pin edge (length ~1e-13, from the rounding of
        // pin-local coordinates) can have its midpoint just outside the pin:
        // find_reg gives -1 and the reference stores -1 + first_reg, i.e. the
        // FSR before this pin (pin_mesh_cyl.cpp:213-217, unchecked). Same here;
        // where that index would be negative the reference reads out of
        // bounds, so that case is an error.
        if (lr >= n_reg || lr + first_reg < 0) throw Error("Ray segment outside of its pin mesh");
        reg.push_back(lr + first_reg);
    }
    return (int)ps.size() - 1;
}

// ---------------------------------------------------------------------------
// core mesh: pins (core/pin.cpp), lattices (lattice.cpp), assemblies
// (assembly.cpp), core (core.cpp), plane (plane.cpp), core mesh (core_mesh.cpp)
// ---------------------------------------------------------------------------
namespace {
Boundary bc_parse(const XmlNode &n, const char *s)
{
    std::string v = n.attr_str(s);
    if (v == "vacuum") return Boundary::VACUUM;
    if (v == "reflect") return Boundary::REFLECT;
    if (v == "prescribed") return Boundary::PRESCRIBED;
    return Boundary::INVALID;
}
}  // namespace

CoreMesh::CoreMesh(const XmlNode &doc, const std::string &base_dir)
{
    const XmlNode *ml = doc.child("material_lib");
    if (!ml) throw Error("No material library specified.");
    mat_lib.load(*ml, base_dir);

    for (const XmlNode *m : doc.children_named("mesh")) {
        auto pm = PinMesh::from_xml(*m);
        int id = pm->id;
        if (pin_meshes.count(id)) throw Error("Duplicate pin mesh ID (" + std::to_string(id) + ")");
        pin_meshes[id] = std::move(pm);
    }
    for (const XmlNode *pn : doc.children_named("pin")) {
        auto pin = std::make_unique<Pin>();
        pin->id = pn->attr_int("id", 0);
        if (pin->id == 0) throw Error("Failed to read pin ID.");
        int mid = pn->attr_int("mesh", 0);
        if (!pin_meshes.count(mid)) throw Error("Invalid pin mesh ID.");
        pin->mesh = pin_meshes[mid].get();
        pin->mat_ids = explode<int>(pn->text);
        if ((int)pin->mat_ids.size() != pin->mesh->n_xsreg)
            throw Error("Wrong number of materials specified in pin definition");
        for (int m : pin->mat_ids) {
            if (!mat_lib.has(m)) throw Error("Invalid material specified in pin");
            if (mat_lib.by_id(m).fissile) {
                pin->is_fuel = true;
                break;
            }
        }
        if (pn->has_attr("fuel")) pin->is_fuel = pn->attr_bool("fuel", false);
        if (pins.count(pin->id)) throw Error("Duplicate pin ID");
        pins[pin->id] = std::move(pin);
    }
    for (const XmlNode *ln : doc.children_named("lattice")) {
        auto lat = std::make_unique<Lattice>();
        lat->id = ln->attr_int("id", 0);
        if (lat->id == 0) throw Error("Trouble reading lattice ID.");
        lat->nx = ln->attr_int("nx", 0);
        lat->ny = ln->attr_int("ny", 0);
        if (lat->nx == 0 || lat->ny == 0) throw Error("Trouble reading lattice dimensions.");
        VecI ids = explode<int>(ln->text);
        if ((int)ids.size() != lat->nx * lat->ny) throw Error("Incorrect number of pin IDs specified for lattice.");
        lat->pins.resize(lat->nx * lat->ny);
        for (int iy = 0; iy < lat->ny; iy++) {
            int row = lat->ny - iy - 1;
            for (int ix = 0; ix < lat->nx; ix++) {
                int pid = ids[iy * lat->nx + ix];
                if (!pins.count(pid)) throw Error("Unknown pin ID in lattice");
                lat->pins[row * lat->nx + ix] = pins[pid].get();
            }
        }
        for (int ix = 0; ix < lat->nx; ix++) {
            double dx = lat->at(ix, 0).mesh->pitch_x;
            lat->hx += dx;
            lat->hx_vec.push_back(dx);
        }
        for (int iy = 0; iy < lat->ny; iy++) {
            double dy = lat->at(0, iy).mesh->pitch_y;
            lat->hy += dy;
            lat->hy_vec.push_back(dy);
        }
        lat->x_vec.push_back(0.0);
        for (int ix = 0; ix < lat->nx; ix++) lat->x_vec.push_back(lat->x_vec[ix] + lat->hx_vec[ix]);
        lat->y_vec.push_back(0.0);
        for (int iy = 0; iy < lat->ny; iy++) lat->y_vec.push_back(lat->y_vec[iy] + lat->hy_vec[iy]);
        for (int iy = 0; iy < lat->ny; iy++)
            for (int ix = 0; ix < lat->nx; ix++)
                if (lat->at(ix, iy).mesh->pitch_x != lat->hx_vec[ix] ||
                    lat->at(ix, iy).mesh->pitch_y != lat->hy_vec[iy])
                    throw Error("Incongruent pin pitches found in lattice.");
        int prev = 0;
        for (size_t i = 0; i < lat->pins.size(); i++) {
            lat->first_reg_pin.push_back(prev);
            prev += lat->pins[i]->mesh->n_reg;
        }
        lat->n_reg = prev;
        if (lattices.count(lat->id)) throw Error("Duplicate lattice ID");
        lattices[lat->id] = std::move(lat);
    }
    // assemblies: only single-plane (2-D) assemblies are meaningful here
    std::map<int, std::pair<const Lattice *, double>> assemblies;
    for (const XmlNode *an : doc.children_named("assembly")) {
        int id = an->attr_int("id", 0);
        if (id == 0) throw Error("Invalid assembly ID.");
        int np = an->attr_int("np", 0);
        if (np == 0) throw Error("Invalid number of planes (nz) when parsing assembly.");
        double h = an->attr_double("hz", 0.0);
        if (const XmlNode *hzn = an->child("hz")) {
            VecF hzs = explode<double>(hzn->text);
            if ((int)hzs.size() != np) throw Error("Wrong number of plane heights provided for assembly");
            h = hzs[0];
        }
        const XmlNode *lats = an->child("lattices");
        if (!lats) throw Error("No lattices specified!");
        auto blocks = explode_braces(lats->text);
        VecI ids;
        for (auto &b : blocks)
            for (int v : b) ids.push_back(v);
        if ((int)ids.size() != np) throw Error("Incorrect number of lattices specified for assembly.");
        for (int v : ids)
            if (v != ids[0]) throw Error("Only 2-D (single unique plane) assemblies are supported");
        if (np != 1) throw Error("Only single-plane (2-D) assemblies are supported by the MoC path");
        if (!lattices.count(ids[0])) throw Error("Unrecognized lattice ID in assembly.");
        assemblies[id] = {lattices[ids[0]].get(), h};
    }
    // core (core.cpp:43-134), single enabled <core>
    const XmlNode *core = nullptr;
    for (const XmlNode *cn : doc.children_named("core")) {
        bool en = cn->has_attr("enabled") ? cn->attr_bool("enabled", true) : true;
        if (en) {
            if (core) throw Error("More than one enabled core specification found.");
            core = cn;
        }
    }
    if (!core) throw Error("No enabled core specifications.");
    nlat_x = core->attr_int("nx", 0);
    nlat_y = core->attr_int("ny", 0);
    if (nlat_x < 1 || nlat_y < 1) throw Error("Invalid core dimensions.");
    bc[NORTH] = bc_parse(*core, "north");
    bc[SOUTH] = bc_parse(*core, "south");
    bc[EAST] = bc_parse(*core, "east");
    bc[WEST] = bc_parse(*core, "west");
    bc[TOP] = bc_parse(*core, "top");
    bc[BOTTOM] = bc_parse(*core, "bottom");
    for (auto b : bc)
        if (b == Boundary::INVALID) throw Error("Not all boundary conditions properly specified.");
    VecI asy = explode<int>(core->text);
    if ((int)asy.size() != nlat_x * nlat_y) throw Error("Wrong number of assemblies specified for core.");
    plane_lattices.assign(nlat_x * nlat_y, nullptr);
    int iasy = 0;
    hz = -1.0;
    for (int iy = 0; iy < nlat_y; iy++) {
        int row = nlat_y - iy - 1;
        for (int ix = 0; ix < nlat_x; ix++) {
            int aid = asy[iasy++];
            if (!assemblies.count(aid)) throw Error("Failed to locate assembly in core specification.");
            plane_lattices[row * nlat_x + ix] = assemblies[aid].first;
            double h = assemblies[aid].second;
            if (hz < 0.0) hz = h;
            else if (!fp_equiv_ulp(h, hz)) throw Error("Assemblies in the core are not compatible.");
        }
    }
    auto at = [&](int ix, int iy) -> const Lattice & { return *plane_lattices[iy * nlat_x + ix]; };
    // Plane (plane.cpp:24-103)
    lat_x.push_back(0.0);
    double prev = 0.0;
    for (int ix = 0; ix < nlat_x; ix++) {
        lat_x.push_back(prev + at(ix, 0).hx);
        prev += at(ix, 0).hx;
    }
    lat_y.push_back(0.0);
    prev = 0.0;
    for (int iy = 0; iy < nlat_y; iy++) {
        lat_y.push_back(prev + at(0, iy).hy);
        prev += at(0, iy).hy;
    }
    for (int iy = 0; iy < nlat_y; iy++)
        for (int ix = 0; ix < nlat_x; ix++) {
            if (!fp_equiv_ulp(at(ix, iy).hx, at(ix, 0).hx) || !fp_equiv_ulp(at(ix, iy).hy, at(0, iy).hy))
                throw Error("Lattices do not have compatible dimensions.");
            if (at(ix, iy).nx != at(ix, 0).nx || at(ix, iy).ny != at(0, iy).ny)
                throw Error("Lattices do not have compatible pin counts.");
        }
    int reg = 0;
    for (const Lattice *l : plane_lattices) {
        first_reg_lattice.push_back(reg);
        reg += l->n_reg;
        for (const Pin *p : l->pins) {
            core_pins.push_back(p);
            if (p->is_fuel) n_fuel++;
        }
    }
    n_reg = reg;
    nx = 0;
    for (int ix = 0; ix < nlat_x; ix++) nx += at(ix, 0).nx;
    ny = 0;
    for (int iy = 0; iy < nlat_y; iy++) ny += at(0, iy).ny;
    hx_core = 0.0;
    for (int ix = 0; ix < nlat_x; ix++) hx_core += at(ix, 0).hx;
    hy_core = 0.0;
    for (int iy = 0; iy < nlat_y; iy++) hy_core += at(0, iy).hy;
    // global pin boundaries (core_mesh.cpp:97-122)
    x_vec.push_back(0.0);
    double hp = 0.0;
    for (int ilx = 0; ilx < nlat_x; ilx++)
        for (double h : at(ilx, 0).hx_vec) {
            x_vec.push_back(h + hp);
            hp += h;
        }
    y_vec.push_back(0.0);
    hp = 0.0;
    for (int ily = 0; ily < nlat_y; ily++)
        for (double h : at(0, ily).hy_vec) {
            y_vec.push_back(h + hp);
            hp += h;
        }
    // Internal pin boundary lines (core_mesh.cpp:124-130). The reference ends
    // the vertical lines at y = hx_core; using max(hx, hy) reproduces it for
    // square cores and still covers taller-than-wide ones.
    double ylen = std::max(hx_core, hy_core);
    for (size_t i = 1; i + 1 < x_vec.size(); i++) lines.push_back(Line(Point2(x_vec[i], 0.0), Point2(x_vec[i], ylen)));
    for (size_t i = 1; i + 1 < y_vec.size(); i++) lines.push_back(Line(Point2(0.0, y_vec[i]), Point2(hx_core, y_vec[i])));
    // FSR offsets of each pin in FSR order
    int r = 0;
    for (const Pin *p : core_pins) {
        pin_first_reg.push_back(r);
        r += p->mesh->n_reg;
    }
}

// Plane::pin_position (plane.cpp:125-152): FSR-order pin -> global x, y
int CoreMesh::pin_position_x(int ipin) const
{
    int ilat = 0;
    for (const Lattice *l : plane_lattices) {
        if (ipin < (int)l->pins.size()) break;
        ipin -= (int)l->pins.size();
        ilat++;
    }
    int lx = ilat % nlat_x;
    int x = 0;
    for (int ix = 0; ix < lx; ix++) x += plane_lattices[ix]->nx;
    return x + ipin % plane_lattices[ilat]->nx;
}

int CoreMesh::pin_position_y(int ipin) const
{
    int ilat = 0;
    for (const Lattice *l : plane_lattices) {
        if (ipin < (int)l->pins.size()) break;
        ipin -= (int)l->pins.size();
        ilat++;
    }
    int ly = ilat / nlat_x;
    int y = 0;
    for (int iy = 0; iy < ly; iy++) y += plane_lattices[iy * nlat_x]->ny;
    return y + ipin / plane_lattices[ilat]->nx;
}

VecF CoreMesh::areas() const
{
    VecF a;
    for (const Pin *p : core_pins)
        for (double v : p->mesh->areas) a.push_back(v);
    return a;
}

// mesh.cpp:99-147 (plane 0 of a one-plane mesh)
int CoreMesh::coarse_surf(int cell, Surface s) const
{
    int ix = cell % nx, iy = cell / nx;
    int nxsurf = nx + 1, nysurf = ny + 1;
    int nxysurf = (nx + 1) * ny + (ny + 1) * nx;
    switch (s) {
    case EAST: return nxsurf * iy + ix + 1 + nx * ny;
    case WEST: return nxsurf * iy + ix + nx * ny;
    case NORTH: return nx * ny + (nx + 1) * ny + (ny + 1) * ix + iy + 1;
    case SOUTH: return nxsurf * ny + nysurf * ix + iy + nx * ny;
    case BOTTOM: return nx * iy + ix;
    case TOP: return nx * ny + nxysurf + nx * iy + ix;
    default: throw Error("coarse_surf: invalid surface");
    }
}

int CoreMesh::coarse_neighbor(int cell, Surface s) const
{
    int ix = cell % nx, iy = cell / nx;
    switch (s) {
    case NORTH: return (iy < ny - 1) ? cell + nx : -1;
    case SOUTH: return (iy > 0) ? cell - nx : -1;
    case EAST: return (ix < nx - 1) ? cell + 1 : -1;
    case WEST: return (ix > 0) ? cell - 1 : -1;
    case TOP:
    case BOTTOM: return -1;
    default: return -5;
    }
}

namespace {
// The first entry of the ascending v that equals p (fp_equiv) or lies above
// it: `on` if it equals p, and i is its index, or the index before it when it
// lies above (v.size() when there is none). The entries equal to p form one
// run around p, so "equal to p or above it" holds on a suffix of v and a
// binary search finds its start.
void locate(const VecF &v, double p, int &i, bool &on)
{
    const auto it = std::partition_point(v.begin(), v.end(), [p](double xi) { return !(fp_equiv(p, xi) || xi > p); });
    i = (int)(it - v.begin());
    on = it != v.end() && fp_equiv(p, *it);
    if (it != v.end() && !on) i--;
}
}  // namespace

// mesh.cpp:390-572
size_t CoreMesh::coarse_norm_point(const Point2 &p, int octant, Surface (&s)[2]) const
{
    int ix, iy;
    bool on_x, on_y;
    locate(x_vec, p.x, ix, on_x);
    locate(y_vec, p.y, iy, on_y);
    if (!on_x && !on_y) return 0;
    if (on_x != on_y) {
        if (on_x) {
            if ((ix == 0) && ((octant == 1) || (octant == 4))) { s[0] = WEST; return 1; }
            if ((ix == nx) && ((octant == 2) || (octant == 3))) { s[0] = EAST; return 1; }
        }
        if (on_y) {
            if ((iy == 0) && ((octant == 1) || (octant == 2))) { s[0] = SOUTH; return 1; }
            if ((iy == ny) && ((octant == 3) || (octant == 4))) { s[0] = NORTH; return 1; }
        }
        if (on_x) {
            if ((octant == 1) || (octant == 4)) s[0] = EAST;
            if ((octant == 2) || (octant == 3)) s[0] = WEST;
        } else {
            if ((octant == 1) || (octant == 2)) s[0] = NORTH;
            if ((octant == 3) || (octant == 4)) s[0] = SOUTH;
        }
        return 1;
    }
    if ((ix > 0) && (ix < nx) && (iy > 0) && (iy < ny)) {
        switch (octant) {
        case 1: s[0] = EAST; s[1] = NORTH; break;
        case 2: s[0] = WEST; s[1] = NORTH; break;
        case 3: s[0] = WEST; s[1] = SOUTH; break;
        case 4: s[0] = EAST; s[1] = SOUTH; break;
        }
        return 2;
    }
    if (ix == 0) {
        switch (octant) {
        case 1: s[0] = WEST; return 1;
        case 2: s[0] = NORTH; s[1] = WEST; return 2;
        case 3: s[0] = SOUTH; s[1] = WEST; return 2;
        case 4: s[0] = WEST; return 1;
        }
    }
    if (ix == nx) {
        switch (octant) {
        case 1: s[0] = NORTH; s[1] = EAST; return 2;
        case 2: s[0] = EAST; return 1;
        case 3: s[0] = EAST; return 1;
        case 4: s[0] = SOUTH; s[1] = EAST; return 2;
        }
    }
    if (iy == 0) {
        switch (octant) {
        case 1: s[0] = SOUTH; return 1;
        case 2: s[0] = SOUTH; return 1;
        case 3: s[0] = WEST; s[1] = SOUTH; return 2;
        case 4: s[0] = EAST; s[1] = SOUTH; return 2;
        }
    }
    if (iy == ny) {
        switch (octant) {
        case 1: s[0] = EAST; s[1] = NORTH; return 2;
        case 2: s[0] = WEST; s[1] = NORTH; return 2;
        case 3: s[0] = NORTH; return 1;
        case 4: s[0] = NORTH; return 1;
        }
    }
    return 0;
}

// mesh.cpp:574-665
int CoreMesh::coarse_boundary_cell(const Point2 &p, int octant) const
{
    int ix, iy;
    bool on_x, on_y;
    locate(x_vec, p.x, ix, on_x);
    locate(y_vec, p.y, iy, on_y);
    if (fp_equiv_abs(p.x, 0.0)) {
        if (octant != 1 && on_y) iy--;
    } else if (fp_equiv(p.x, hx_core)) {
        ix--;
        if (octant != 2 && on_y) iy--;
    } else if (fp_equiv_abs(p.y, 0.0)) {
        if (octant != 1 && on_x) ix--;
    } else if (fp_equiv(p.y, hy_core)) {
        iy--;
        if (octant == 3 && on_x) ix--;
    } else {
        throw Error("coarse_boundary_cell: point not on the domain boundary");
    }
    int cell = iy * nx + ix;
    if (cell < 0 || cell >= nx * ny || ix < 0 || ix >= nx || iy < 0 || iy >= ny)
        throw Error("coarse_boundary_cell: cell out of range");
    return cell;
}

// CoreMesh::get_pinmesh -> Plane::get_pinmesh -> Lattice::get_pinmesh with the
// default Direction (positive x and y cosines) used by Ray::Ray.
const PinMesh *CoreMesh::get_pinmesh(Point2 &p, int &first_reg) const
{
    auto lb = [](const VecF &v, double x) {
        return (int)(std::lower_bound(v.begin(), v.end(), x, fuzzy_lt) - v.begin());
    };
    int ix = lb(lat_x, p.x);
    if (ix < (int)lat_x.size() && fp_equiv(p.x, lat_x[ix])) ix = ix + 1;
    ix--;
    ix = std::min(nlat_x - 1, std::max(0, ix));
    int iy = lb(lat_y, p.y);
    if (iy < (int)lat_y.size() && fp_equiv(p.y, lat_y[iy])) iy = iy + 1;
    iy--;
    iy = std::min(nlat_y - 1, std::max(0, iy));
    int ilat = nlat_x * iy + ix;
    p.x -= lat_x[ix];
    p.y -= lat_y[iy];
    first_reg += first_reg_lattice[ilat];
    const Lattice &lat = *plane_lattices[ilat];
    int px = lb(lat.x_vec, p.x);
    if (px < (int)lat.x_vec.size() && fp_equiv_abs(p.x, lat.x_vec[px])) px = px + 1;
    px--;
    px = std::min(lat.nx - 1, std::max(0, px));
    p.x = 0.5 * (lat.x_vec[px + 1] + lat.x_vec[px]);
    int py = lb(lat.y_vec, p.y);
    if (py < (int)lat.y_vec.size() && fp_equiv_abs(p.y, lat.y_vec[py])) py = py + 1;
    py--;
    py = std::min(lat.ny - 1, std::max(0, py));
    p.y = 0.5 * (lat.y_vec[py + 1] + lat.y_vec[py]);
    first_reg += lat.first_reg_pin[py * lat.nx + px];
    p.x += lat_x[ix];
    p.y += lat_y[iy];
    return lat.at(px, py).mesh;
}

void CoreMesh::trace(std::vector<Point2> &ps) const
{
    Line l(ps[0], ps[1]);
    for (const auto &li : lines) {
        Point2 q;
        if (intersect_lines(li, l, q) == 1) ps.push_back(q);
    }
    std::sort(ps.begin(), ps.end());
    ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
    snap_corners(ps);
}

// A ray that passes within rounding distance of a pin corner crosses the
// corner's x line and y line at two points a few ULPs apart (one of them may
// be the ray's end on the domain boundary). Point2 equality (200 ULPs per
// coordinate) may keep both while coarse_norm_point classifies one of them
// as the corner (both lines) and the other as a crossing of one line: the
// coarse data then count one surface twice and the pin-cell currents no
// longer balance. Crossings of the pin lines closer than kCornerSnap can only
// be such a pair, so they become one point exactly on the corner (a ray end
// stays the ray end), which coarse_norm_point takes as one corner crossing.
// A point that it takes as a corner anyway is also moved onto the corner, so
// that no sliver of a segment is traced in the neighbouring pin.
void CoreMesh::snap_corners(std::vector<Point2> &ps) const
{
    auto nearest = [](const VecF &v, double x) {
        size_t i = (size_t)(std::lower_bound(v.begin(), v.end(), x) - v.begin());
        if (i == v.size() || (i > 0 && x - v[i - 1] < v[i] - x)) i--;
        return v[i];
    };
    auto corner = [&](const Point2 &p) { return Point2(nearest(x_vec, p.x), nearest(y_vec, p.y)); };
    const size_t n = ps.size();
    std::vector<Point2> out;
    out.reserve(n);
    for (size_t k = 0; k < n; k++) {
        if (k + 1 < n && std::hypot(ps[k + 1].x - ps[k].x, ps[k + 1].y - ps[k].y) < kCornerSnap) {
            out.push_back(corner(ps[k]));
            k++;  // the pair is one crossing
            continue;
        }
        const Point2 c = corner(ps[k]);
        out.push_back((fp_equiv(ps[k].x, c.x) && fp_equiv(ps[k].y, c.y)) ? c : ps[k]);
    }
    ps.swap(out);
}

// ---------------------------------------------------------------------------
// rays: sweepers/moc/ray.cpp:52-176
// ---------------------------------------------------------------------------
Ray::Ray(Point2 a, Point2 b, int bc0, int bc1, const CoreMesh &mesh) : p1(a), p2(b)
{
    bc[0] = bc0;
    bc[1] = bc1;
    std::vector<Point2> ps{a, b};
    mesh.trace(ps);
    std::vector<Point2> cps = ps;
    VecI cm_nseg;
    Point2 p_prev = cps.front();
    for (size_t k = 1; k < ps.size(); k++) {
        Point2 pin_p = Midpoint(ps[k], p_prev);
        int first_reg = 0;
        const PinMesh *pm = mesh.get_pinmesh(pin_p, first_reg);
        int ns = pm->trace(p_prev - pin_p, ps[k] - pin_p, first_reg, seg_len, seg_index);
        cm_nseg.push_back(ns);
        p_prev = ps[k];
    }
    // coarse ray data
    Surface s[2];
    int octant = (b.x > a.x) ? 1 : 2;
    std::vector<int> surfs_fw, nsegs_fw, surfs_bw, nsegs_bw;
    {
        cm_cell_fw = mesh.coarse_boundary_cell(cps.front(), octant);
        size_t ns = mesh.coarse_norm_point(cps.front(), octant, s);
        if (ns != 1) throw Error("Ray: first coarse crossing must be one surface");
        cm_surf_fw = mesh.coarse_surf(cm_cell_fw, s[0]);
        for (size_t k = 0; k < cm_nseg.size(); k++) {
            ns = mesh.coarse_norm_point(cps[k + 1], octant, s);
            for (size_t i = 0; i < ns; i++) surfs_fw.push_back(s[i]);
            nsegs_fw.push_back(cm_nseg[k]);
            if (ns > 1) nsegs_fw.push_back(0);
        }
    }
    std::reverse(cm_nseg.begin(), cm_nseg.end());
    octant = (octant == 1) ? 3 : 4;
    {
        cm_cell_bw = mesh.coarse_boundary_cell(cps.back(), octant);
        size_t ns = mesh.coarse_norm_point(cps.back(), octant, s);
        if (ns != 1) throw Error("Ray: first coarse crossing must be one surface");
        cm_surf_bw = mesh.coarse_surf(cm_cell_bw, s[0]);
        for (size_t k = 0; k < cm_nseg.size(); k++) {
            ns = mesh.coarse_norm_point(cps[cps.size() - 2 - k], octant, s);
            for (size_t i = 0; i < ns; i++) surfs_bw.push_back(s[i]);
            nsegs_bw.push_back(cm_nseg[k]);
            if (ns > 1) nsegs_bw.push_back(0);
        }
    }
    size_t stp = std::min(nsegs_fw.size(), nsegs_bw.size());
    for (size_t i = 0; i < stp; i++) cm_data.push_back({surfs_fw[i], surfs_bw[i], nsegs_fw[i], nsegs_bw[i]});
    if (nsegs_fw.size() > nsegs_bw.size()) cm_data.push_back({surfs_fw.back(), INVALID, 0, 0});
    if (nsegs_bw.size() > nsegs_fw.size()) cm_data.push_back({INVALID, surfs_bw.back(), 0, 0});
}

int host_threads()
{
    for (const char *name : {"MOCC_HOST_THREADS", "OMP_NUM_THREADS"}) {
        const char *v = std::getenv(name);
        if (v && std::atoi(v) > 0) return std::min(64, std::atoi(v));
    }
    const int hw = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(64, hw));
}

void parallel_for(int n, const std::function<void(int)> &f)
{
    const int nt = std::min(n, host_threads());
    if (nt <= 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&]() {
        for (int i = next++; i < n && !failed.load(); i = next++) {
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                if (!err) err = std::current_exception();
                failed = true;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (err) std::rethrow_exception(err);
}

// util/rational_approximation.hpp: binary search of the Farey sequence
std::pair<int, int> rational_approximation(double target, double tol, int max_d)
{
    if (max_d == 0) max_d = std::numeric_limits<int>::max();
    int a = 0, b = 1, c = 1, d = 1;
    if (std::abs(double(a) / double(b) - target) < tol) return {a, b};
    if (std::abs(double(c) / double(d) - target) < tol) return {c, d};
    while ((b <= max_d) && (d <= max_d)) {
        double mediant = double(a + c) / double(b + d);
        if (std::abs(target - mediant) < tol) {
            if (b + d <= max_d) return {a + c, b + d};
            if (d > b) return {c, d};
            return {a, b};
        } else if (target > mediant) {
            a = a + c;
            b = b + d;
        } else {
            c = a + c;
            d = b + d;
        }
    }
    if (b > max_d) return {c, d};
    return {a, b};
}

// ray_data.cpp:491-527 (RATIONAL modularization, the default)
std::pair<int, int> RayData::modularize_angle(const Angle &ang, double hx, double hy, double spacing) const
{
    double m = (ang.alpha < kPI / 4.0) ? std::tan(ang.alpha) : std::tan(kPI / 2.0 - ang.alpha);
    auto nxy = rational_approximation(m, 0.002, 0);
    int Nx = nxy.first, Ny = nxy.second;
    if (Nx == 0) throw Error("Azimuthal angle too close to an axis to modularize");
    double alpha = std::atan((Nx * hy) / (Ny * hx));
    int scale = (int)std::ceil((hx / Nx * std::abs(std::sin(alpha))) / spacing);
    Nx *= scale;
    Ny *= scale;
    if (ang.alpha > kPI / 4.0) std::swap(Nx, Ny);
    return {Nx, Ny};
}

// ray_data.cpp:67-343
RayData::RayData(const XmlNode &in, const AngularQuadrature &quad, const CoreMesh &mesh) : ang_quad(quad)
{
    double opt = in.attr_double("spacing", -1.0);
    if (opt <= 0.0) throw Error("Failed to read valid ray spacing.");
    std::string vc = in.attr_str("volume_correction", "flat");
    std::transform(vc.begin(), vc.end(), vc.begin(), ::tolower);
    if (vc == "flat") correction = VolumeCorrection::FLAT;
    else if (vc == "angle") correction = VolumeCorrection::ANGLE;
    else if (vc == "none") correction = VolumeCorrection::NONE;
    else throw Error("Unrecognized volume correction option in <rays>");
    std::string mod = in.attr_str("modularity", "core");
    if (mod != "core") throw Error("Only core-modular ray tracing is supported");
    std::string mz = in.attr_str("modularization", "rational_fraction");
    if (mz != "rational_fraction") throw Error("Only rational_fraction modularization is supported");

    double hx = mesh.hx_core, hy = mesh.hy_core;
    int nd = ang_quad.ndir_oct();
    for (int iang = 0; iang < nd; iang++) {
        Angle ang = ang_quad[iang];
        auto N = modularize_angle(ang, hx, hy, opt);
        Nx.push_back(N.first);
        Ny.push_back(N.second);
        double new_alpha = std::atan(hy * N.first / (hx * N.second));
        ang.modify_alpha(new_alpha);
        ang_quad.modify_angle(iang, ang);
        spacing.push_back(std::cos(ang_quad[iang].alpha) * hy / N.second);
    }
    ang_quad.update_weights();
    for (int iang = 0; iang < nd * 3; iang++) {
        Nx.push_back(Nx[iang]);
        Ny.push_back(Ny[iang]);
        spacing.push_back(spacing[iang]);
    }

    // Trace one ray set per azimuth. Product quadratures share the azimuth
    // across n_polar consecutive angles; other quadratures have one angle per
    // azimuth.
    int n_pol = (ang_quad.type() == QuadType::LS || ang_quad.type() == QuadType::USER) ? 1 : ang_quad.n_polar();
    int n_azi = 2 * nd / n_pol;
    // The azimuths are independent: each is traced on its own thread into
    // its own slot, so the rays are the same for any thread count.
    Point2 lo(0.0, 0.0), hi(hx, hy);
    rays.resize(n_azi);
    VecI azi_max(n_azi, 0);
    parallel_for(n_azi, [&](int a) {
        int iang = a * n_pol;
        const Angle &ang = ang_quad[iang];
        int nx_ = Nx[iang], ny_ = Ny[iang];
        double space = spacing[iang];
        double sx = std::abs(space / std::sin(ang.alpha));
        double sy = std::abs(space / std::cos(ang.alpha));
        std::vector<Ray> &rs = rays[a];
        rs.reserve(nx_ + ny_);
        auto exit_bc = [&](const Point2 &p2) -> int {
            if (fp_equiv(p2.x, hx)) return (int)(p2.y / sy);
            if (fp_equiv(p2.y, hy)) return (int)(p2.x / sx) + ny_;
            if (fp_equiv(p2.x, 0.0)) return (int)(p2.y / sy);
            throw Error("Something has gone horribly wrong in the ray trace.");
        };
        for (int ir = 0; ir < ny_; ir++) {
            Point2 p1;
            p1.x = (ang.ox > 0.0) ? 0.0 : hx;
            p1.y = (0.5 + ir) * sy;
            p1.ok = true;
            Point2 p2 = box_intersect(lo, hi, p1, ang.alpha);
            if (!p2.ok) throw Error("Ray does not exit the core box");
            rs.emplace_back(p1, p2, ir, exit_bc(p2), mesh);
            azi_max[a] = std::max(azi_max[a], rs.back().nseg());
        }
        for (int ir = 0; ir < nx_; ir++) {
            Point2 p1((0.5 + ir) * sx, 0.0);
            Point2 p2 = box_intersect(lo, hi, p1, ang.alpha);
            if (!p2.ok) throw Error("Ray does not exit the core box");
            rs.emplace_back(p1, p2, ir + ny_, exit_bc(p2), mesh);
            azi_max[a] = std::max(azi_max[a], rs.back().nseg());
        }
    });
    for (int m : azi_max) max_seg = std::max(max_seg, m);
    correct_volume(mesh);
}

// ray_data.cpp:345-489
void RayData::correct_volume(const CoreMesh &mesh)
{
    VecF true_vol = mesh.areas();
    int nreg = (int)true_vol.size();
    int n_pol = (int)(2 * ang_quad.ndir_oct() / rays.size());
    if (correction == VolumeCorrection::FLAT) {
        // each azimuth is corrected on its own
        parallel_for((int)rays.size(), [&](int a) {
            double space = spacing[(size_t)a * n_pol];
            VecF fsr_vol(nreg, 0.0);
            for (const auto &r : rays[a])
                for (int k = 0; k < r.nseg(); k++) fsr_vol[r.seg_index[k]] += r.seg_len[k] * space;
            VecF cf(nreg);
            for (int i = 0; i < nreg; i++) cf[i] = true_vol[i] / fsr_vol[i];
            for (auto &r : rays[a])
                for (int k = 0; k < r.nseg(); k++) r.seg_len[k] = r.seg_len[k] * cf[r.seg_index[k]];
        });
    } else if (correction == VolumeCorrection::ANGLE) {
        VecF fsr_vol(nreg, 0.0);
        for (size_t a = 0; a < rays.size(); a++)
            for (int p = 0; p < n_pol; p++) {
                int iang = (int)a * n_pol + p;
                double space = spacing[iang];
                double wgt = ang_quad[iang].weight * 0.5;
                for (const auto &r : rays[a])
                    for (int k = 0; k < r.nseg(); k++) fsr_vol[r.seg_index[k]] += r.seg_len[k] * space * wgt;
            }
        for (int i = 0; i < nreg; i++) fsr_vol[i] = true_vol[i] / fsr_vol[i];
        for (auto &rs : rays)
            for (auto &r : rs)
                for (int k = 0; k < r.nseg(); k++) r.seg_len[k] = r.seg_len[k] * fsr_vol[r.seg_index[k]];
    }
}

}  // namespace mocc_amd
