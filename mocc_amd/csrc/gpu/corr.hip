// The 2-D/3-D correction worker on the device (corr.hpp): the FP64 re-sweep
// of one group that restates cmdo::CurrentCorrections::post_ray /
// post_angle / calculate_corrections (correction_worker.hpp:114-262,
// correction_worker.cpp:33-160), and MoCSweeper::update_incoming_flux
// (moc_sweeper.cpp:255-291, update_incoming_generic moc_sweeper.hpp:198-233).
#include "corr.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <stdexcept>
#include <string>

namespace mocc_amd {
namespace gpu {

namespace {

constexpr double kPi = 3.1415926535897932;
constexpr double kFourPi = 4.0 * kPi;
enum { EAST = 0, NORTH = 1, WEST = 2, SOUTH = 3, TOP = 4, BOTTOM = 5, INVALID = 7 };

#define CORR_CHECK(expr)                                                                          \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

// Mesh::coarse_surf / coarse_neighbor of one plane (src/core/mesh.cpp:99-147).
__host__ __device__ inline int coarse_surf(int nx, int ny, int cell, int s)
{
    const int ix = cell % nx, iy = cell / nx;
    switch (s) {
    case EAST: return (nx + 1) * iy + ix + 1 + nx * ny;
    case WEST: return (nx + 1) * iy + ix + nx * ny;
    case NORTH: return nx * ny + (nx + 1) * ny + (ny + 1) * ix + iy + 1;
    case SOUTH: return (nx + 1) * ny + (ny + 1) * ix + iy + nx * ny;
    default: return -1;
    }
}
inline int coarse_neighbor(int nx, int ny, int cell, int s)
{
    const int ix = cell % nx, iy = cell / nx;
    switch (s) {
    case NORTH: return (iy < ny - 1) ? cell + nx : -1;
    case SOUTH: return (iy > 0) ? cell - nx : -1;
    case EAST: return (ix < nx - 1) ? cell + 1 : -1;
    case WEST: return (ix > 0) ? cell - 1 : -1;
    default: return -1;
    }
}

// Exponential_Linear<10000> on [-10, 0] in FP64 (src/core/exponential.hpp:56-112):
// d[i] = exp(-10 + i / 1000), linear between nodes; exp itself outside.
__device__ inline double exp_linear(double v)
{
    const double mn = -10.0, space = 10.0 / 10000.0, rspace = 1.0 / space;
    if (v < mn || v > 0.0) return exp(v);
    int i = (int)((v - mn) * rspace);
    if (i > 9999) i = 9999;
    const double d0 = exp(mn + i * space), d1 = exp(mn + (i + 1) * space);
    v -= space * i + mn;
    return d0 + (d1 - d0) * v * rspace;
}

struct WalkArgs {
    int R, A, NP, G, g, phase, gs;
    const int32_t *seg_fsr;
    const double *seg_len;
    const int64_t *ray_seg_off;
    const int32_t *ray_info;
    const int32_t *ent_nseg;
    const AziInfo *azi;
    const double *ang;
    const int32_t *fsr_mat;
    const double *mat_tr, *q;
    double *bc, *ent, *ray_psi0;
};

// One (ray, direction, polar) of the correction sweep: the ray's angular flux
// in FP64 from the snapshot of the incoming flux, with CurrentCorrections'
// per-visit sums (fluxvol = t q + dpsi / Sigma_t, its Sigma_t-weighted twin,
// and the forward track length) and the flux at each crossing. Under the
// Gauss-Seidel update the octant-1 rays (phase 0) leave their exit flux in
// the incoming slots of their mirror angles, which the octant-2 rays (phase 1)
// read, as the band sweep does.
__global__ void k_corr_walk(WalkArgs w)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w.R * 2 * w.NP) return;
    const int r = t / (2 * w.NP), dir = (t / w.NP) % 2, p = t % w.NP;
    const int32_t *ri = w.ray_info + 8 * (size_t)r;
    const int bc0 = ri[0], bc1 = ri[1], a = ri[2], flags = ri[3], c0 = ri[4], c1 = ri[5];
    if (w.gs && (w.phase == 0) != (a < w.A / 2)) return;
    const AziInfo az = w.azi[a];
    const int in_off = dir ? az.in_off[1] : az.in_off[0];
    double psi = w.bc[in_off + p * az.stride + (dir ? bc1 : bc0)];
    w.ray_psi0[((size_t)r * 2 + dir) * w.NP + p] = psi;
    const double rs = w.ang[((size_t)a * w.NP + p) * 8 + 0];
    int64_t k = dir ? w.ray_seg_off[r + 1] - 1 : w.ray_seg_off[r];
    const int G = w.G, g = w.g;
    for (int c = c0; c < c1; c++) {
        const int n = w.ent_nseg[2 * (size_t)c + dir];
        double vol = 0.0, sig = 0.0, vn = 0.0;
        for (int i = 0; i < n; i++) {
            const int f = w.seg_fsr[k];
            const double st = w.mat_tr[w.fsr_mat[f] * G + g];
            const double qq = w.q[f];
            const double l = w.seg_len[k];
            const double tl = rs * l;
            const double e = 1.0 - exp_linear(-st * l * rs);
            const double pn = psi - (psi - qq) * e;
            // forward: (psi1[k] - psi1[k+1]) / xstr; backward: e (psi2[k+1] - q) / xstr
            const double fv = dir ? tl * qq + e * (psi - qq) / st : tl * qq + (psi - pn) / st;
            vol += fv;
            sig += st * fv;
            if (!dir) vn += tl;
            psi = pn;
            k += dir ? -1 : 1;
        }
        double *o = w.ent + (((size_t)c * 2 + dir) * w.NP + p) * 4;
        o[0] = vol;
        o[1] = sig;
        o[2] = vn;
        o[3] = psi;
    }
    const int vac = (flags >> (2 + dir)) & 1;
    if (w.gs && w.phase == 0 && !vac) {
        const int ny = (flags >> dir) & 1;
        const int out = ny ? az.out_off[dir][1] : az.out_off[dir][0];
        w.bc[out + p * az.stride + (dir ? bc0 : bc1)] = psi;
    }
}

struct CellArgs {
    int A, NP, nd, nx, ny, n_cell, G, g;
    const int32_t *vis_ptr, *vis_item, *crs_ptr, *crs_item;
    const double *ang, *dx, *dy, *htr, *ent, *ray_psi0;
    double *alpha, *beta, *res_part;
};

// The cell face a direction enters or leaves through (correction_worker.cpp:
// 46-67): q = 0, 1 the x faces XL, XR, q = 2, 3 the y faces YL, YR; the moc
// angles all run +y, their reverses -y, and ox decides east against west.
__device__ inline int face_side(int dir, int q, bool ox_pos)
{
    if (q >= 2) return ((dir == 0) == (q == 2)) ? SOUTH : NORTH;
    return (((q == 0) == (dir == 0)) == ox_pos) ? WEST : EAST;
}

// Sum of the crossing fluxes of (azimuth a, surface s) in direction dir,
// polar p, in list order.
__device__ inline double crossing_sum(const CellArgs &c, int n_surf, int a, int s, int dir, int p)
{
    double v = 0.0;
    if (s < 0) return v;
    const size_t key = (size_t)a * n_surf + s;
    for (int k = c.crs_ptr[key]; k < c.crs_ptr[key + 1]; k++) {
        const int it = c.crs_item[k];
        if (((it >> 1) & 1) != dir) continue;
        const int idx = it >> 2;
        v += (it & 1) ? c.ray_psi0[((size_t)idx * 2 + dir) * c.NP + p]
                      : c.ent[(((size_t)idx * 2 + dir) * c.NP + p) * 4 + 3];
    }
    return v;
}

// post_angle's normalization and calculate_corrections for (azimuth, polar,
// cell): both directions (the angle of octant 1 or 2 and its reverse).
__global__ void k_corr_cells(CellArgs c, int n_surf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.A * c.NP * c.n_cell) return;
    const int cell = t % c.n_cell, ap = t / c.n_cell;
    const int a = ap / c.NP, p = ap % c.NP;
    const int iang = a * c.NP + p, nang = 4 * c.nd, iang2 = (iang + 2 * c.nd) % nang;
    const double *an = c.ang + (size_t)ap * 8;
    const double alpha_ang = an[1], spacing = an[2], ox = an[7];
    double vol[2] = {0.0, 0.0}, sig[2] = {0.0, 0.0}, vn = 0.0;
    const size_t key = (size_t)a * c.n_cell + cell;
    for (int k = c.vis_ptr[key]; k < c.vis_ptr[key + 1]; k++) {
        const int it = c.vis_item[k];
        const int dir = it & 1, e = it >> 1;
        const double *o = c.ent + (((size_t)e * 2 + dir) * c.NP + p) * 4;
        vol[dir] += o[0];
        sig[dir] += o[1];
        if (!dir) vn += o[2];
    }
    const double area0 = fabs(spacing / cos(alpha_ang)), area1 = fabs(spacing / sin(alpha_ang));
    const double area_x = area0 / c.dx[cell % c.nx], area_y = area1 / c.dy[cell / c.nx];
    const double xs = c.htr[(size_t)cell * c.G + c.g];
    double res[3] = {0.0, 0.0, 0.0};
    for (int dir = 0; dir < 2; dir++) {
        const double sg = sig[dir] / vol[dir];
        const double va = vol[dir] / vn;
        const bool ox_pos = ox > 0.0;
        auto face = [&](int q) {
            return crossing_sum(c, n_surf, a, coarse_surf(c.nx, c.ny, cell, face_side(dir, q, ox_pos)), dir, p);
        };
        const double psi_xl = face(0) * area_x, psi_xr = face(1) * area_x;
        const double psi_yl = face(2) * area_y, psi_yr = face(3) * area_y;
        const double ax = va / (psi_xl + psi_xr);
        const double ay = va / (psi_yl + psi_yr);
        const double b = sg / xs;
        const int ia = dir ? iang2 : iang;
        const size_t ib = ((size_t)c.g * nang + ia) * c.n_cell + cell;
        double &al_x = c.alpha[ib * 2 + 0], &al_y = c.alpha[ib * 2 + 1], &be = c.beta[ib];
        res[0] += (ax - al_x) * (ax - al_x);
        res[1] += (ay - al_y) * (ay - al_y);
        res[2] += (b - be) * (b - be);
        al_x = ax;
        al_y = ay;
        be = b;
    }
    for (int q = 0; q < 3; q++) c.res_part[(size_t)t * 3 + q] = res[q];
}

// The residual triple of the group: each sum in a fixed order (strided
// partial sums, then a fixed tree), then the square root (residual()).
__global__ __launch_bounds__(256) void k_corr_resid(const double *part, int n, double *out)
{
    __shared__ double sh[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 256)
        for (int q = 0; q < 3; q++) s[q] += part[(size_t)i * 3 + q];
    for (int q = 0; q < 3; q++) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = sqrt(sh[threadIdx.x][0]);
}

// The coarse current and surface flux of the worker (post_ray: + for the
// forward crossings, - for the backward ones, in both the current and, as
// the reference's worker has it, the surface flux), over the surface's area
// (moc::Current::post_sweep). Radial surfaces n0 <= s < n1 of group g.
__global__ void k_corr_currents(CellArgs c, int n_surf, int n0, int n1, const double *area, double *cur,
                                double *sfl)
{
    const int s = n0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n1) return;
    const int nxy = c.nx * c.ny;
    const int normal = (s - nxy) < (c.nx + 1) * c.ny ? 0 : 1;
    double j = 0.0, f = 0.0;
    for (int a = 0; a < c.A; a++) {
        const size_t key = (size_t)a * n_surf + s;
        for (int k = c.crs_ptr[key]; k < c.crs_ptr[key + 1]; k++) {
            const int it = c.crs_item[k];
            const int dir = (it >> 1) & 1, idx = it >> 2;
            for (int p = 0; p < c.NP; p++) {
                const double psi = (it & 1) ? c.ray_psi0[((size_t)idx * 2 + dir) * c.NP + p]
                                            : c.ent[(((size_t)idx * 2 + dir) * c.NP + p) * 4 + 3];
                const double *an = c.ang + ((size_t)a * c.NP + p) * 8;
                const double cw = an[3 + normal], fw = an[5 + normal];
                if (dir) {
                    j -= psi * cw;
                    f -= psi * fw;
                } else {
                    j += psi * cw;
                    f += psi * fw;
                }
            }
        }
    }
    cur[(size_t)s * c.G + c.g] = j / area[s];
    sfl[(size_t)s * c.G + c.g] = f / area[s];
}

__global__ void k_corr_q(int F, int G, int g, const int32_t *fsr_mat, const double *mat_tr, const double *mat_ss,
                         const double *src, const double *phi, double *q)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int m = fsr_mat[f];
    const double tr = mat_tr[m * G + g];
    q[f] = (src[(size_t)f * G + g] + phi[(size_t)f * G + g] * mat_ss[m * G + g]) * (1.0 / (tr * kFourPi));
}

__global__ void k_corr_bc(const float *psi, int n, double *bc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bc[i] = (double)psi[i];
}

__global__ void k_fill_d(double *a, double v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// update_incoming_generic: item (group, angle of octants 1-2, ray); the
// reference passes the ray's first coarse cell as the partial-current index.
__global__ void k_update_incoming(int R, int NP, int G, int bcpg, uint64_t parity, const int32_t *ray_info,
                                  const AziInfo *azi, const double *part, const double *part_old, int has_old,
                                  float *psi)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)G * R * NP) return;
    const int p = (int)(t % NP), r = (int)((t / NP) % R), g = (int)(t / ((size_t)NP * R));
    const int32_t *ri = ray_info + 8 * (size_t)r;
    const AziInfo az = azi[ri[2]];
    const int par = (int)((parity >> g) & 1ull);
    float *in = psi + ((size_t)g * 2 + par) * bcpg;
    for (int dir = 0; dir < 2; dir++) {
        const int is = dir ? ri[7] : ri[6];
        const size_t k = ((size_t)is * G + g) * 2;
        float &v = in[(dir ? az.in_off[1] : az.in_off[0]) + p * az.stride + (dir ? ri[1] : ri[0])];
        if (has_old) {
            const double pn = 2.0 * (part[k] + part[k + 1]);
            const double po = 2.0 * (part_old[k] + part_old[k + 1]);
            if (po > 0.0) v = (float)((double)v * (pn / po));
        } else {
            v = (float)(2.0 * (1.0 / kFourPi) * (part[k] + part[k + 1]));
        }
    }
}

}  // namespace

CorrSolver::~CorrSolver() { free_all(); }

void CorrSolver::free_all()
{
    for (void *p : owned_) (void)hipFree(p);
    owned_.clear();
}

size_t CorrSolver::bytes() const
{
    if (!full_) return (size_t)R_ * 8 * 4;
    return (size_t)S_ * 12 + (size_t)n_cm_ * (2 * 4 + 2 * NP_ * 4 * 8) + (size_t)R_ * (8 * 4 + 2 * NP_ * 8) +
           (size_t)G_ * 4 * nd_ * n_cell_ * 3 * 8;
}

void CorrSolver::setup(const mocgpu_problem &p, const SweepPlan &plan, hipStream_t s, bool full)
{
    free_all();
    R_ = p.n_ray;
    A_ = p.n_azi;
    NP_ = p.n_polar;
    nd_ = A_ * NP_ / 2;
    nx_ = p.nx;
    ny_ = p.ny;
    n_cell_ = nx_ * ny_;
    n_surf_ = (nx_ + 1) * ny_ + (ny_ + 1) * nx_ + 2 * nx_ * ny_;
    S_ = p.n_seg;
    G_ = p.n_group;
    bcpg_ = plan.bcpg;
    n_cm_ = p.ray_cm_off[R_];
    auto up = [&](auto *&dst, const auto &v) {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        void *q = nullptr;
        CORR_CHECK(hipMalloc(&q, std::max<size_t>(v.size(), 1) * sizeof(T)));
        owned_.push_back(q);
        if (!v.empty()) CORR_CHECK(hipMemcpyAsync(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        dst = (std::remove_reference_t<decltype(dst)>)q;
    };
    auto alloc = [&](double *&dst, size_t n) {
        void *q = nullptr;
        CORR_CHECK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(double)));
        owned_.push_back(q);
        dst = (double *)q;
    };
    full_ = full;
    if (full) {
        up(d_seg_fsr_, std::vector<int32_t>(p.seg_fsr, p.seg_fsr + S_));
        up(d_seg_len_, std::vector<double>(p.seg_len, p.seg_len + S_));
        up(d_ray_seg_off_, std::vector<int64_t>(p.ray_seg_off, p.ray_seg_off + R_ + 1));
    }

    // rays: boundary indices, azimuth, exit flags (as the band plan forms
    // them), coarse entries and first cells; with the worker also the
    // per-entry segment counts and the visit and crossing lists of each
    // (azimuth, cell / surface). Without it (update_incoming_flux alone) the
    // ray table is all: every CMFD handle builds it, so it stays O(rays).
    std::vector<int32_t> info((size_t)R_ * 8), nseg(full ? (size_t)2 * n_cm_ : 0, 0);
    std::vector<std::vector<int32_t>> vis(full ? (size_t)A_ * n_cell_ : 0), crs(full ? (size_t)A_ * n_surf_ : 0);
    for (int a = 0; a < A_; a++) {
        const int Ny = p.azi_ny[a];
        const int oct = a < A_ / 2 ? 0 : 1;
        for (int r = p.azi_ray_off[a]; r < p.azi_ray_off[a + 1]; r++) {
            int32_t *ri = &info[(size_t)r * 8];
            ri[0] = p.ray_bc[2 * r];
            ri[1] = p.ray_bc[2 * r + 1];
            ri[2] = a;
            const int fw_y = ri[1] >= Ny ? 1 : 0, bw_y = ri[0] >= Ny ? 1 : 0;
            const int fw_face = fw_y ? NORTH : (oct == 0 ? EAST : WEST);
            const int bw_face = bw_y ? SOUTH : (oct == 0 ? WEST : EAST);
            ri[3] = fw_y | (bw_y << 1) | ((p.bc[fw_face] == MOCGPU_BC_VACUUM) << 2) |
                    ((p.bc[bw_face] == MOCGPU_BC_VACUUM) << 3);
            ri[4] = p.ray_cm_off[r];
            ri[5] = p.ray_cm_off[r + 1];
            ri[6] = p.ray_cm_cell[2 * r];
            ri[7] = p.ray_cm_cell[2 * r + 1];
            if (!full) continue;
            for (int dir = 0; dir < 2; dir++) {
                const int s0 = p.ray_cm_surf[2 * r + dir];
                if (s0 >= 0 && s0 < n_surf_) crs[(size_t)a * n_surf_ + s0].push_back((r << 2) | (dir << 1) | 1);
                int cell = p.ray_cm_cell[2 * r + dir];
                for (int c = ri[4]; c < ri[5]; c++) {
                    const int32_t *crd = &p.cm_data[4 * (size_t)c];
                    const int surf = crd[dir];
                    if (surf != INVALID) {
                        nseg[2 * (size_t)c + dir] = crd[2 + dir];
                        if (cell >= 0) {
                            vis[(size_t)a * n_cell_ + cell].push_back((c << 1) | dir);
                            const int sg = coarse_surf(nx_, ny_, cell, surf);
                            if (sg >= 0) crs[(size_t)a * n_surf_ + sg].push_back((c << 2) | (dir << 1));
                        }
                    }
                    cell = cell >= 0 ? coarse_neighbor(nx_, ny_, cell, surf) : -1;
                }
            }
        }
    }
    auto csr = [](const std::vector<std::vector<int32_t>> &lists, std::vector<int32_t> &ptr,
                  std::vector<int32_t> &item) {
        ptr.assign(lists.size() + 1, 0);
        for (size_t i = 0; i < lists.size(); i++) ptr[i + 1] = ptr[i] + (int32_t)lists[i].size();
        item.reserve(ptr.back());
        for (const auto &l : lists) item.insert(item.end(), l.begin(), l.end());
    };
    up(d_ray_info_, info);
    up(d_azi_, plan.azi);
    if (!full) return;
    std::vector<int32_t> vp, vi, cp, ci;
    csr(vis, vp, vi);
    csr(crs, cp, ci);
    up(d_ent_nseg_, nseg);
    up(d_vis_ptr_, vp);
    up(d_vis_item_, vi);
    up(d_crs_ptr_, cp);
    up(d_crs_item_, ci);

    // moc::Current::set_angle weights (moc_current_worker.hpp:182-200) and
    // the angle's data per (azimuth, polar)
    std::vector<double> ang((size_t)A_ * NP_ * 8);
    for (int a = 0; a < A_; a++)
        for (int q = 0; q < NP_; q++) {
            const int ia = a * NP_ + q;
            const double w = p.ang_weight[ia] * kPi, sp = p.azi_spacing[a], al = p.ang_alpha[ia], dz = p.hz;
            double *o = &ang[(size_t)ia * 8];
            o[0] = p.ang_rsin_theta[ia];
            o[1] = al;
            o[2] = sp;
            o[3] = w * p.ang_ox[ia] * sp / std::abs(std::cos(al)) * dz;
            o[4] = w * p.ang_oy[ia] * sp / std::abs(std::sin(al)) * dz;
            o[5] = w * sp / std::abs(std::cos(al)) * dz;
            o[6] = w * sp / std::abs(std::sin(al)) * dz;
            o[7] = p.ang_ox[ia];
        }
    up(d_ang_, ang);
    std::vector<double> dx(nx_), dy(ny_);
    for (int i = 0; i < nx_; i++) dx[i] = p.x_vec[i + 1] - p.x_vec[i];
    for (int i = 0; i < ny_; i++) dy[i] = p.y_vec[i + 1] - p.y_vec[i];
    up(d_dx_, dx);
    up(d_dy_, dy);
    up(d_area_, plan.surf_area);

    alloc(d_q_, (size_t)p.n_fsr);
    alloc(d_bc_, (size_t)bcpg_);
    alloc(d_ent_, (size_t)n_cm_ * 2 * NP_ * 4);
    alloc(d_ray_psi0_, (size_t)R_ * 2 * NP_);
    alloc(d_res_part_, (size_t)A_ * NP_ * n_cell_ * 3);
    const size_t na = (size_t)G_ * 4 * nd_ * n_cell_;
    alloc(d_alpha_, 2 * na);
    alloc(d_beta_, na);
    alloc(d_resid_, (size_t)G_ * 3);
    // CorrectionData starts at the diamond-difference values
    hipLaunchKernelGGL(k_fill_d, dim3(nblk(2 * na)), dim3(256), 0, s, d_alpha_, 0.5, 2 * na);
    hipLaunchKernelGGL(k_fill_d, dim3(nblk(na)), dim3(256), 0, s, d_beta_, 1.0, na);
    hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)G_ * 3)), dim3(256), 0, s, d_resid_, 0.0, (size_t)G_ * 3);
    CORR_CHECK(hipGetLastError());
}

void CorrSolver::before_sweep(const CorrView &v, hipStream_t s)
{
    hipLaunchKernelGGL(k_corr_q, dim3(nblk(v.F)), dim3(256), 0, s, v.F, v.G, v.g, v.fsr_mat, v.mat_tr, v.mat_ss,
                       v.src, v.phi, d_q_);
    const int par = (int)(((uint64_t)v.parity >> v.g) & 1ull);
    hipLaunchKernelGGL(k_corr_bc, dim3(nblk(bcpg_)), dim3(256), 0, s, v.psi + ((size_t)v.g * 2 + par) * bcpg_,
                       bcpg_, d_bc_);
    CORR_CHECK(hipGetLastError());
}

void CorrSolver::after_sweep(const CorrView &v, hipStream_t s)
{
    WalkArgs w{};
    w.R = R_;
    w.A = A_;
    w.NP = NP_;
    w.G = v.G;
    w.g = v.g;
    w.gs = v.gs;
    w.seg_fsr = d_seg_fsr_;
    w.seg_len = d_seg_len_;
    w.ray_seg_off = d_ray_seg_off_;
    w.ray_info = d_ray_info_;
    w.ent_nseg = d_ent_nseg_;
    w.azi = d_azi_;
    w.ang = d_ang_;
    w.fsr_mat = v.fsr_mat;
    w.mat_tr = v.mat_tr;
    w.q = d_q_;
    w.bc = d_bc_;
    w.ent = d_ent_;
    w.ray_psi0 = d_ray_psi0_;
    const int nt = R_ * 2 * NP_;
    for (int ph = 0; ph < (v.gs ? 2 : 1); ph++) {
        w.phase = ph;
        hipLaunchKernelGGL(k_corr_walk, dim3(nblk(nt, 64)), dim3(64), 0, s, w);
    }
    CellArgs c{};
    c.A = A_;
    c.NP = NP_;
    c.nd = nd_;
    c.nx = nx_;
    c.ny = ny_;
    c.n_cell = n_cell_;
    c.G = v.G;
    c.g = v.g;
    c.vis_ptr = d_vis_ptr_;
    c.vis_item = d_vis_item_;
    c.crs_ptr = d_crs_ptr_;
    c.crs_item = d_crs_item_;
    c.ang = d_ang_;
    c.dx = d_dx_;
    c.dy = d_dy_;
    c.htr = v.htr;
    c.ent = d_ent_;
    c.ray_psi0 = d_ray_psi0_;
    c.alpha = d_alpha_;
    c.beta = d_beta_;
    c.res_part = d_res_part_;
    const int ncells = A_ * NP_ * n_cell_;
    hipLaunchKernelGGL(k_corr_cells, dim3(nblk(ncells)), dim3(256), 0, s, c, n_surf_);
    hipLaunchKernelGGL(k_corr_resid, dim3(1), dim3(256), 0, s, d_res_part_, ncells, d_resid_ + 3 * (size_t)v.g);
    const int n0 = nx_ * ny_, n1 = n0 + (nx_ + 1) * ny_ + (ny_ + 1) * nx_;
    hipLaunchKernelGGL(k_corr_currents, dim3(nblk(n1 - n0)), dim3(256), 0, s, c, n_surf_, n0, n1, d_area_, v.cur,
                       v.sfl);
    CORR_CHECK(hipGetLastError());
}

void CorrSolver::corrections(hipStream_t s, double *alpha, double *beta) const
{
    const size_t na = (size_t)G_ * 4 * nd_ * n_cell_;
    if (alpha) CORR_CHECK(hipMemcpyAsync(alpha, d_alpha_, 2 * na * sizeof(double), hipMemcpyDeviceToHost, s));
    if (beta) CORR_CHECK(hipMemcpyAsync(beta, d_beta_, na * sizeof(double), hipMemcpyDeviceToHost, s));
    CORR_CHECK(hipStreamSynchronize(s));
}

void CorrSolver::residual(hipStream_t s, int g, double *out3) const
{
    CORR_CHECK(hipMemcpyAsync(out3, d_resid_ + 3 * (size_t)g, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    CORR_CHECK(hipStreamSynchronize(s));
}

void CorrSolver::update_incoming(float *psi, uint64_t parity, int G, const double *part, const double *part_old,
                                 bool has_old, hipStream_t s) const
{
    const size_t n = (size_t)G * R_ * NP_;
    hipLaunchKernelGGL(k_update_incoming, dim3(nblk(n)), dim3(256), 0, s, R_, NP_, G, bcpg_, parity, d_ray_info_,
                       d_azi_, part, part_old, has_old ? 1 : 0, psi);
    CORR_CHECK(hipGetLastError());
}

}  // namespace gpu
}  // namespace mocc_amd
