// FP64 CPU oracle for the MoC sweep, the source update and pin-cell CMFD.
//
// TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load this library as the checker. The product
// (mocc_amd/, libmocgpu.so) never links, loads or calls it.
//
// Each function restates the reference algorithm it cites, with the same loop
// order and OpenMP structure (thread-private FSR tallies reduced in a critical
// section, schedule(static,1) over the rays of one angle), so that it also
// serves as the timed CPU baseline.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr double PI = 3.1415926535897932;
constexpr double FPI = 4.0 * PI;
enum { EAST = 0, NORTH = 1, WEST = 2, SOUTH = 3, TOP = 4, BOTTOM = 5, INVALID = 7 };
enum { BC_VACUUM = 0, BC_REFLECT = 1, BC_PRESCRIBED = 4 };

thread_local std::string g_err;

// Exponential_Linear<N> (src/core/exponential.hpp:56-112). The index is
// clamped to N-1 so that v == max does not read d[N+1].
struct ExpTable {
    int n;
    double min, max, space, rspace;
    std::vector<double> d;
    ExpTable(int n_ = 10000, double mn = -10.0, double mx = 0.0)
        : n(n_), min(mn), max(mx), space((mx - mn) / (double)n_), rspace(1.0 / space), d(n_ + 1)
    {
        for (int i = 0; i <= n; i++) d[i] = std::exp(min + i * space);
    }
    double exp(double v) const
    {
        if (v < min || v > max) return std::exp(v);
        int i = (int)((v - min) * rspace);
        if (i > n - 1) i = n - 1;
        v -= space * i + min;
        return d[i] + (d[i + 1] - d[i]) * v * rspace;
    }
};

struct Oracle {
    mocgpu_problem p{};
    oracle_options o{};
    // deep copies of the problem arrays
    std::vector<double> ang_weight, ang_sin, ang_rsin, ang_ox, ang_oy, ang_alpha, azi_spacing;
    std::vector<int32_t> azi_ray_off, azi_nx, azi_ny, ray_bc, seg_fsr, ray_cm_cell, ray_cm_surf,
        ray_cm_off, cm_data, fsr_mat, pin_fsr_off, pin_cell;
    std::vector<int64_t> ray_seg_off;
    std::vector<double> seg_len, fsr_vol, fsr_area, xs_tr, xs_nf, xs_f, xs_ch, xs_sc, x_vec, y_vec, pin_area;

    int F = 0, G = 0, nd = 0, A = 0, NP = 0, max_seg = 0;
    ExpTable table;
    std::vector<double> flux, flux_old, xstr, source_1g, q, fs;
    bool has_fs = false;
    // boundary condition storage (boundary_condition.cpp:33-78)
    int bcpg = 0;
    std::vector<int> off;        // [4nd][3]
    std::vector<int> fsize;      // [4nd][3]
    std::vector<double> bc_in;   // [G * bcpg]
    std::vector<double> bc_out;  // [bcpg]
    // coarse data (coarse_data.cpp)
    int nx = 0, ny = 0, n_surf = 0, n_cell = 0;
    std::vector<double> cur, sfl, part, part_old, cflux;
    bool has_radial = false, has_old_partial = false;
    // CMFD state
    int n_solve = 0;
    std::vector<double> h_tr, h_nf, h_f, h_ch, h_rm, h_sc;  // homogenized per cell
    std::vector<double> d_tilde, d_hat, s_tilde, s_hat;    // [n_surf*G]
    std::vector<std::vector<int>> m_rowptr, m_col;
    std::vector<std::vector<double>> m_val;
    int last_cmfd_iters = 0;
    // cmdo::CorrectionData (correction_data.hpp:45-66): alpha [G][4nd][cell][2]
    // (initially 0.5), beta [G][4nd][cell] (initially 1); the last residual
    // triple of each group's correction sweep
    std::vector<double> alpha, beta, corr_resid;

    int angle_face_size(int iang, int face) const { return fsize[iang * 3 + face]; }
};

template <typename T> void copy_vec(std::vector<T> &dst, const T *src, size_t n)
{
    dst.assign(src, src + n);
}

// AngularQuadrature::reverse / reflect (angular_quadrature.hpp:120-167)
inline int reverse_ang(const Oracle &o, int iang) { return (iang + o.nd * 2) % (o.nd * 4); }
inline int reflect_ang(const Oracle &o, int iang, int normal)
{
    static const int refl[3][8] = {{1, 0, 3, 2, 5, 4, 7, 6}, {3, 2, 1, 0, 7, 6, 5, 4}, {4, 5, 6, 7, 0, 1, 2, 3}};
    int ioct = iang / o.nd;
    return iang + (refl[normal][ioct] - ioct) * o.nd;
}
// Direction::upwind_surface for the angle of index iang (octants 1-4)
inline int upwind_surface(const Oracle &o, int iang, int normal)
{
    int ioct = iang / o.nd;  // 0..3
    bool ox_pos = (ioct == 0 || ioct == 3);
    bool oy_pos = (ioct == 0 || ioct == 1);
    if (normal == 0) return ox_pos ? WEST : EAST;
    return oy_pos ? SOUTH : NORTH;
}

// ---- coarse mesh helpers (src/core/mesh.cpp:99-147, mesh.hpp) -------------
int coarse_surf(const Oracle &o, int cell, int s)
{
    int nx = o.nx, ny = o.ny;
    int ix = cell % nx, iy = cell / nx;
    int nxysurf = (nx + 1) * ny + (ny + 1) * nx;
    switch (s) {
    case EAST: return (nx + 1) * iy + ix + 1 + nx * ny;
    case WEST: return (nx + 1) * iy + ix + nx * ny;
    case NORTH: return nx * ny + (nx + 1) * ny + (ny + 1) * ix + iy + 1;
    case SOUTH: return (nx + 1) * ny + (ny + 1) * ix + iy + nx * ny;
    case BOTTOM: return nx * iy + ix;
    case TOP: return nx * ny + nxysurf + nx * iy + ix;
    }
    throw std::runtime_error("coarse_surf: bad surface");
}
int coarse_neighbor(const Oracle &o, int cell, int s)
{
    int ix = cell % o.nx, iy = cell / o.nx;
    switch (s) {
    case NORTH: return (iy < o.ny - 1) ? cell + o.nx : -1;
    case SOUTH: return (iy > 0) ? cell - o.nx : -1;
    case EAST: return (ix < o.nx - 1) ? cell + 1 : -1;
    case WEST: return (ix > 0) ? cell - 1 : -1;
    case TOP:
    case BOTTOM: return -1;
    }
    return -5;
}
int surface_normal(const Oracle &o, int surf)
{
    int nsurfz = o.nx * o.ny + (o.nx + 1) * o.ny + (o.ny + 1) * o.nx;
    int s = surf % nsurfz;
    if (s < o.nx * o.ny) return 2;
    if (s < o.nx * o.ny + (o.nx + 1) * o.ny) return 0;
    return 1;
}
int surface_to_normal(int s) { return (s == EAST || s == WEST) ? 0 : ((s == NORTH || s == SOUTH) ? 1 : 2); }
double dx_of(const Oracle &o, int ix) { return o.x_vec[ix + 1] - o.x_vec[ix]; }
double dy_of(const Oracle &o, int iy) { return o.y_vec[iy + 1] - o.y_vec[iy]; }
// Mesh::coarse_area(surf)
double coarse_area_surf(const Oracle &o, int surf)
{
    int n = surface_normal(o, surf);
    int nsp = o.nx * o.ny + (o.nx + 1) * o.ny + (o.ny + 1) * o.nx;
    int iz = surf / nsp;
    int s = surf - iz * nsp;
    if (n == 0) return dy_of(o, (s - o.nx * o.ny) / (o.nx + 1)) * o.p.hz;
    if (n == 1) {
        s -= o.nx * o.ny + (o.nx + 1) * o.ny;
        return dx_of(o, s / (o.ny + 1)) * o.p.hz;
    }
    return dx_of(o, s % o.nx) * dy_of(o, s / o.nx);
}
// Mesh::coarse_area(cell, surf)
double coarse_area_cell(const Oracle &o, int cell, int s)
{
    int ix = cell % o.nx, iy = cell / o.nx;
    switch (s) {
    case EAST:
    case WEST: return dy_of(o, iy) * o.p.hz;
    case NORTH:
    case SOUTH: return dx_of(o, ix) * o.p.hz;
    default: return dx_of(o, ix) * dy_of(o, iy);
    }
}
// Mesh::coarse_neigh_cells
std::pair<int, int> coarse_neigh_cells(const Oracle &o, int surf)
{
    int nx = o.nx, ny = o.ny;
    int nsp = nx * ny + (nx + 1) * ny + (ny + 1) * nx;
    int n = surface_normal(o, surf);
    int iz = surf / nsp;
    int s = surf - iz * nsp;
    int nz = 1;
    std::pair<int, int> c;
    if (n == 0) {
        int iy = (s - nx * ny) / (nx + 1), ix = (s - nx * ny) % (nx + 1);
        c.first = ix > 0 ? iz * nx * ny + iy * nx + ix - 1 : -1;
        c.second = ix < nx ? iz * nx * ny + iy * nx + ix : -1;
    } else if (n == 1) {
        s -= nx * ny + (nx + 1) * ny;
        int iy = s % (ny + 1), ix = s / (ny + 1);
        c.first = iy > 0 ? iz * nx * ny + (iy - 1) * nx + ix : -1;
        c.second = iy < ny ? iz * nx * ny + iy * nx + ix : -1;
    } else {
        int iy = s / nx, ix = s % nx;
        c.first = iz > 0 ? (iz - 1) * nx * ny + iy * nx + ix : -1;
        c.second = iz < nz ? iz * nx * ny + iy * nx + ix : -1;
    }
    return c;
}

// ---- boundary conditions (src/core/boundary_condition.cpp) -----------------
void bc_initialize_scalar(Oracle &o, double val)
{
    std::fill(o.bc_in.begin(), o.bc_in.end(), 0.0);
    for (int g = 0; g < o.G; g++)
        for (int iang = 0; iang < 4 * o.nd; iang++)
            for (int n = 0; n < 2; n++) {
                int s = upwind_surface(o, iang, n);
                int kind = o.p.bc[s];
                if (kind == BC_REFLECT) {
                    double *face = &o.bc_in[g * o.bcpg + o.off[iang * 3 + n]];
                    for (int i = 0; i < o.angle_face_size(iang, n); i++) face[i] = val;
                } else if (kind == BC_PRESCRIBED) {
                    double *face = &o.bc_in[g * o.bcpg + o.off[iang * 3 + n]];
                    for (int i = 0; i < o.angle_face_size(iang, n); i++) face[i] = 1.0 / FPI * 2.0;
                }
            }
}
// BoundaryCondition::update(group, angle, out)
void bc_update(Oracle &o, int g, int angle)
{
    for (int n = 0; n < 3; n++) {
        int size = o.angle_face_size(angle, n);
        if (size == 0) break;
        int iang_in = reflect_ang(o, angle, n);
        int off_in = g * o.bcpg + o.off[iang_in * 3 + n];
        int off_out = o.off[angle * 3 + n];
        int kind = o.p.bc[upwind_surface(o, iang_in, n)];
        if (kind == BC_VACUUM) {
            for (int i = 0; i < size; i++) o.bc_in[off_in + i] = 0.0;
        } else if (kind == BC_REFLECT) {
            for (int i = 0; i < size; i++) o.bc_in[off_in + i] = o.bc_out[off_out + i];
        } else if (kind != BC_PRESCRIBED) {
            throw std::runtime_error("Unsupported boundary condition type");
        }
    }
}

// ---- source (src/core/source.cpp:41-112, source_isotropic.cpp:21-57) --------
void source_initialize(Oracle &o) { std::fill(o.source_1g.begin(), o.source_1g.end(), 0.0); }
void source_fission(Oracle &o, int g)
{
    // loop over XS regions then their FSRs; region order = material order
    for (int f = 0; f < o.F; f++) {
        int m = o.fsr_mat[f];
        o.source_1g[f] += o.xs_ch[m * o.G + g] * o.fs[f];
    }
}
void source_in_scatter(Oracle &o, int g, const std::vector<double> &phi)
{
    for (int f = 0; f < o.F; f++) {
        int m = o.fsr_mat[f];
        const double *row = &o.xs_sc[(size_t)m * o.G * o.G + (size_t)g * o.G];
        for (int gg = 0; gg < o.G; gg++) {
            if (gg == g) continue;
            o.source_1g[f] += row[gg] * phi[(size_t)f * o.G + gg];
        }
    }
}
void source_self_scatter(Oracle &o, int g)
{
    for (int f = 0; f < o.F; f++) {
        int m = o.fsr_mat[f];
        double xssc = o.xs_sc[(size_t)m * o.G * o.G + (size_t)g * o.G + g];
        double r_fpi_tr = 1.0 / (o.xs_tr[m * o.G + g] * FPI);
        o.q[f] = (o.source_1g[f] + o.flux[(size_t)f * o.G + g] * xssc) * r_fpi_tr;
    }
}

void calculate_corrections(Oracle &o, int iang, int g, const std::vector<double> &surf_sum,
                           const std::vector<double> &vol_sum, const std::vector<double> &sigt_sum,
                           double resid[3]);
void homogenize(Oracle &o);

// ---- the sweep: moc_sweeper_kernel.inc.hpp:36-180 with moc::Current --------
// worker 0: NoCurrent, 1: moc::Current, 2: cmdo::CurrentCorrections
// (correction_worker.hpp:36-292), which also tallies the per-angle cell and
// surface sums of the 2-D/3-D correction factors.
void sweep1g(Oracle &o, int g, bool with_current, int worker = -1)
{
    if (worker < 0) worker = with_current ? 1 : 0;
    with_current = worker != 0;
    std::vector<double> flux_1g(o.F, 0.0);
    const int n_surf_plane = o.nx * o.ny + (o.nx + 1) * o.ny + (o.ny + 1) * o.nx;
    const int n_cell = o.nx * o.ny;
    // CurrentCorrections' per-angle buffers (correction_worker.hpp:50-56)
    std::vector<double> surf_sum, vol_sum, vol_norm, sigt_sum;
    double resid[3] = {0.0, 0.0, 0.0};
    if (worker == 2) {
        surf_sum.assign(2 * n_surf_plane, 0.0);
        vol_sum.assign(2 * n_cell, 0.0);
        vol_norm.assign(n_cell, 0.0);
        sigt_sum.assign(2 * n_cell, 0.0);
    }
#pragma omp parallel default(shared) num_threads(o.o.n_threads > 0 ? o.o.n_threads : omp_get_max_threads())
    {
        std::vector<double> e_tau(o.max_seg), psi1(o.max_seg + 1), psi2(o.max_seg + 1);
        std::vector<double> t_flux(o.F, 0.0);
        for (int iang = 0; iang < 2 * o.nd; iang++) {
            const int a = iang / o.NP;
            const double *qbar = o.q.data();
            const int iang1 = iang, iang2 = reverse_ang(o, iang);
            const double *bc_in_1 = &o.bc_in[(size_t)g * o.bcpg + o.off[iang1 * 3]];
            double *bc_out_1 = &o.bc_out[o.off[iang1 * 3]];
            const double *bc_in_2 = &o.bc_in[(size_t)g * o.bcpg + o.off[iang2 * 3]];
            double *bc_out_2 = &o.bc_out[o.off[iang2 * 3]];
            const double spacing = o.azi_spacing[a];
            // moc::Current::set_angle (moc_current_worker.hpp:182-200)
            double cw[2], fw[2];
            {
                double w = o.ang_weight[iang] * PI;
                double dz = o.p.hz;
                double al = o.ang_alpha[iang];
                cw[0] = w * o.ang_ox[iang] * spacing / std::abs(std::cos(al)) * dz;
                cw[1] = w * o.ang_oy[iang] * spacing / std::abs(std::sin(al)) * dz;
                fw[0] = w * spacing / std::abs(std::cos(al)) * dz;
                fw[1] = w * spacing / std::abs(std::sin(al)) * dz;
            }
            const double stheta = o.ang_sin[iang];
            const double rstheta = o.ang_rsin[iang];
            const double wt_v_st = o.ang_weight[iang] * spacing * o.p.hz * stheta * PI;
            const int r0 = o.azi_ray_off[a], r1 = o.azi_ray_off[a + 1];
            if (worker == 2) {
                // CurrentCorrections::set_angle: zero the per-angle sums
#pragma omp single
                {
                    std::fill(surf_sum.begin(), surf_sum.end(), 0.0);
                    std::fill(vol_sum.begin(), vol_sum.end(), 0.0);
                    std::fill(vol_norm.begin(), vol_norm.end(), 0.0);
                    std::fill(sigt_sum.begin(), sigt_sum.end(), 0.0);
                }
            }
#pragma omp for schedule(static, 1)
            for (int ir = r0; ir < r1; ir++) {
                const int64_t s0 = o.ray_seg_off[ir];
                const int nseg = (int)(o.ray_seg_off[ir + 1] - s0);
                const int bc1 = o.ray_bc[2 * ir], bc2 = o.ray_bc[2 * ir + 1];
                for (int k = 0; k < nseg; k++) {
                    int ireg = o.seg_fsr[s0 + k];
                    double v = -o.xstr[ireg] * o.seg_len[s0 + k] * rstheta;
                    e_tau[k] = 1.0 - (o.o.exact_exp ? std::exp(v) : o.table.exp(v));
                }
                psi1[0] = bc_in_1[bc1];
                for (int k = 0; k < nseg; k++) {
                    int ireg = o.seg_fsr[s0 + k];
                    double psi_diff = (psi1[k] - qbar[ireg]) * e_tau[k];
                    psi1[k + 1] = psi1[k] - psi_diff;
                    t_flux[ireg] += psi_diff * wt_v_st;
                }
                bc_out_1[bc2] = psi1[nseg];
                psi2[nseg] = bc_in_2[bc2];
                for (int k = nseg - 1; k >= 0; k--) {
                    int ireg = o.seg_fsr[s0 + k];
                    double psi_diff = (psi2[k + 1] - qbar[ireg]) * e_tau[k];
                    psi2[k] = psi2[k + 1] - psi_diff;
                    t_flux[ireg] += psi_diff * wt_v_st;
                }
                bc_out_2[bc1] = psi2[0];
                if (worker == 2) {
                    // CurrentCorrections::post_ray (correction_worker.hpp:114-221)
#pragma omp critical
                    {
                        auto C = [&](int s) -> double & { return o.cur[(size_t)s * o.G + g]; };
                        auto S = [&](int s) -> double & { return o.sfl[(size_t)s * o.G + g]; };
                        int cell_fw = o.ray_cm_cell[2 * ir], cell_bw = o.ray_cm_cell[2 * ir + 1];
                        int surf_fw = o.ray_cm_surf[2 * ir], surf_bw = o.ray_cm_surf[2 * ir + 1];
                        int iseg_fw = 0, iseg_bw = nseg;
                        int norm_fw = surface_normal(o, surf_fw);
                        int norm_bw = surface_normal(o, surf_bw);
                        C(surf_fw) += psi1[iseg_fw] * cw[norm_fw];
                        C(surf_bw) -= psi2[iseg_bw] * cw[norm_bw];
                        S(surf_fw) += psi1[iseg_fw] * fw[norm_fw];
                        S(surf_bw) -= psi2[iseg_bw] * fw[norm_bw];  // as the reference's worker does
                        surf_sum[surf_fw * 2 + 0] += psi1[iseg_fw];
                        surf_sum[surf_bw * 2 + 1] += psi2[iseg_bw];
                        for (int c = o.ray_cm_off[ir]; c < o.ray_cm_off[ir + 1]; c++) {
                            const int *crd = &o.cm_data[4 * c];
                            if (crd[0] != INVALID) {
                                for (int i = 0; i < crd[2]; i++) {
                                    const int ireg = o.seg_fsr[s0 + iseg_fw];
                                    const double xs = o.xstr[ireg];
                                    const double t = rstheta * o.seg_len[s0 + iseg_fw];
                                    const double fluxvol = t * qbar[ireg] + (psi1[iseg_fw] - psi1[iseg_fw + 1]) / xs;
                                    vol_sum[cell_fw * 2 + 0] += fluxvol;
                                    vol_norm[cell_fw] += t;
                                    sigt_sum[cell_fw * 2 + 0] += xs * fluxvol;
                                    iseg_fw++;
                                }
                                norm_fw = surface_to_normal(crd[0]);
                                surf_fw = coarse_surf(o, cell_fw, crd[0]);
                                C(surf_fw) += psi1[iseg_fw] * cw[norm_fw];
                                S(surf_fw) += psi1[iseg_fw] * fw[norm_fw];
                                surf_sum[surf_fw * 2 + 0] += psi1[iseg_fw];
                            }
                            if (crd[1] != INVALID) {
                                for (int i = 0; i < crd[3]; i++) {
                                    iseg_bw--;
                                    const int ireg = o.seg_fsr[s0 + iseg_bw];
                                    const double xs = o.xstr[ireg];
                                    const double t = rstheta * o.seg_len[s0 + iseg_bw];
                                    const double fluxvol =
                                        t * qbar[ireg] + e_tau[iseg_bw] * (psi2[iseg_bw + 1] - qbar[ireg]) / xs;
                                    vol_sum[cell_bw * 2 + 1] += fluxvol;
                                    sigt_sum[cell_bw * 2 + 1] += xs * fluxvol;
                                }
                                norm_bw = surface_to_normal(crd[1]);
                                surf_bw = coarse_surf(o, cell_bw, crd[1]);
                                C(surf_bw) -= psi2[iseg_bw] * cw[norm_bw];
                                S(surf_bw) -= psi2[iseg_bw] * fw[norm_bw];
                                surf_sum[surf_bw * 2 + 1] += psi2[iseg_bw];
                            }
                            cell_fw = coarse_neighbor(o, cell_fw, crd[0]);
                            cell_bw = coarse_neighbor(o, cell_bw, crd[1]);
                        }
                    }
                } else if (with_current) {
                    // moc::Current::post_ray (moc_current_worker.hpp:202-264)
#pragma omp critical
                    {
                        double *current = &o.cur[0];
                        double *surface_flux = &o.sfl[0];
                        auto C = [&](int s) -> double & { return current[(size_t)s * o.G + g]; };
                        auto S = [&](int s) -> double & { return surface_flux[(size_t)s * o.G + g]; };
                        int cell_fw = o.ray_cm_cell[2 * ir], cell_bw = o.ray_cm_cell[2 * ir + 1];
                        int surf_fw = o.ray_cm_surf[2 * ir], surf_bw = o.ray_cm_surf[2 * ir + 1];
                        int iseg_fw = 0, iseg_bw = nseg;
                        int norm_fw = surface_normal(o, surf_fw);
                        int norm_bw = surface_normal(o, surf_bw);
                        C(surf_fw) += psi1[iseg_fw] * cw[norm_fw];
                        C(surf_bw) -= psi2[iseg_bw] * cw[norm_bw];
                        S(surf_fw) += psi1[iseg_fw] * fw[norm_fw];
                        S(surf_bw) += psi2[iseg_bw] * fw[norm_bw];
                        for (int c = o.ray_cm_off[ir]; c < o.ray_cm_off[ir + 1]; c++) {
                            const int *crd = &o.cm_data[4 * c];
                            if (crd[0] != INVALID) {
                                iseg_fw += crd[2];
                                norm_fw = surface_to_normal(crd[0]);
                                surf_fw = coarse_surf(o, cell_fw, crd[0]);
                                C(surf_fw) += psi1[iseg_fw] * cw[norm_fw];
                                S(surf_fw) += psi1[iseg_fw] * fw[norm_fw];
                            }
                            if (crd[1] != INVALID) {
                                iseg_bw -= crd[3];
                                norm_bw = surface_to_normal(crd[1]);
                                surf_bw = coarse_surf(o, cell_bw, crd[1]);
                                C(surf_bw) -= psi2[iseg_bw] * cw[norm_bw];
                                S(surf_bw) += psi2[iseg_bw] * fw[norm_bw];
                            }
                            cell_fw = coarse_neighbor(o, cell_fw, crd[0]);
                            cell_bw = coarse_neighbor(o, cell_bw, crd[1]);
                        }
                    }
                }
            }
            if (worker == 2) {
                // CurrentCorrections::post_angle (correction_worker.hpp:239-262)
#pragma omp single
                {
                    for (int i = 0; i < n_cell; i++) {
                        sigt_sum[2 * i + 0] /= vol_sum[2 * i + 0];
                        sigt_sum[2 * i + 1] /= vol_sum[2 * i + 1];
                        vol_sum[2 * i + 0] /= vol_norm[i];
                        vol_sum[2 * i + 1] /= vol_norm[i];
                    }
                    calculate_corrections(o, iang, g, surf_sum, vol_sum, sigt_sum, resid);
                }
            }
            if (o.o.gs_boundary) {
#pragma omp single
                {
                    bc_update(o, g, iang1);
                    bc_update(o, g, iang2);
                }
            }
        }
        if (!o.o.gs_boundary) {
#pragma omp single
            {
                for (int iang = 0; iang < 4 * o.nd; iang++) bc_update(o, g, iang);
            }
        }
#pragma omp barrier
#pragma omp critical
        {
            for (int i = 0; i < o.F; i++) flux_1g[i] += t_flux[i];
        }
#pragma omp barrier
#pragma omp single
        {
            for (int i = 0; i < o.F; i++)
                o.flux[(size_t)i * o.G + g] = flux_1g[i] / (o.xstr[i] * o.fsr_vol[i]) + o.q[i] * FPI;
            if (with_current) {
                // moc::Current::post_sweep: normalize by surface area
                for (int s = o.nx * o.ny; s < n_surf_plane; s++) {
                    double area = coarse_area_surf(o, s);
                    o.cur[(size_t)s * o.G + g] /= area;
                    o.sfl[(size_t)s * o.G + g] /= area;
                }
            }
        }
    }
    if (worker == 2)
        for (int k = 0; k < 3; k++) o.corr_resid[(size_t)g * 3 + k] = std::sqrt(resid[k]);
}

// CurrentCorrections::calculate_corrections (correction_worker.cpp:33-160)
// for the angle iang of octants 1-2 (forward) and its reverse (backward):
// alpha_x / alpha_y = cell-average angular flux over the sum of the two
// x- (y-) face average fluxes, beta = the flux-weighted Sigma_t of the cell
// over its homogenized Sigma_t (the pin-cell XS mesh, h_tr).
void calculate_corrections(Oracle &o, int iang, int g, const std::vector<double> &surf_sum,
                           const std::vector<double> &vol_sum, const std::vector<double> &sigt_sum,
                           double resid[3])
{
    const int n_cell = o.nx * o.ny, nang = 4 * o.nd;
    const int iang1 = iang, iang2 = reverse_ang(o, iang);
    const double ox = o.ang_ox[iang];
    int surfs[2][4];  // [dir][XL, XR, YL, YR]: the moc angles all run +y
    surfs[0][2] = SOUTH;
    surfs[0][3] = NORTH;
    surfs[1][2] = NORTH;
    surfs[1][3] = SOUTH;
    if (ox > 0.0) {
        surfs[0][0] = WEST;
        surfs[0][1] = EAST;
        surfs[1][0] = EAST;
        surfs[1][1] = WEST;
    } else {
        surfs[0][0] = EAST;
        surfs[0][1] = WEST;
        surfs[1][0] = WEST;
        surfs[1][1] = EAST;
    }
    const double spacing = o.azi_spacing[iang / o.NP];
    const double alpha_ang = o.ang_alpha[iang];
    const double area[2] = {std::abs(spacing / std::cos(alpha_ang)), std::abs(spacing / std::sin(alpha_ang))};
    for (int ic = 0; ic < n_cell; ic++) {
        const double area_x = area[0] / dx_of(o, ic % o.nx);
        const double area_y = area[1] / dy_of(o, ic / o.nx);
        const double xs = o.h_tr[(size_t)ic * o.G + g];
        for (int d = 0; d < 2; d++) {
            const int ia = d == 0 ? iang1 : iang2;
            auto ss = [&](int k) { return surf_sum[coarse_surf(o, ic, surfs[d][k]) * 2 + d]; };
            const double psi_xl = ss(0) * area_x, psi_xr = ss(1) * area_x;
            const double psi_yl = ss(2) * area_y, psi_yr = ss(3) * area_y;
            const double ax = vol_sum[ic * 2 + d] / (psi_xl + psi_xr);
            const double ay = vol_sum[ic * 2 + d] / (psi_yl + psi_yr);
            const double b = sigt_sum[ic * 2 + d] / xs;
            double &al_x = o.alpha[(((size_t)g * nang + ia) * n_cell + ic) * 2 + 0];
            double &al_y = o.alpha[(((size_t)g * nang + ia) * n_cell + ic) * 2 + 1];
            double &be = o.beta[((size_t)g * nang + ia) * n_cell + ic];
            resid[0] += (ax - al_x) * (ax - al_x);
            resid[1] += (ay - al_y) * (ay - al_y);
            resid[2] += (b - be) * (b - be);
            al_x = ax;
            al_y = ay;
            be = b;
        }
    }
}

// MoCSweeper::sweep (moc_sweeper.cpp:189-225)
void mocsweep(Oracle &o, int g, bool with_current)
{
    for (int f = 0; f < o.F; f++) o.xstr[f] = o.xs_tr[o.fsr_mat[f] * o.G + g];
    const int n_surf_plane = o.nx * o.ny + (o.nx + 1) * o.ny + (o.ny + 1) * o.nx;
    for (int inner = 0; inner < o.o.n_inner; inner++) {
        source_self_scatter(o, g);
        if (inner == o.o.n_inner - 1 && with_current) {
            // CoarseData::zero_data_radial (coarse_data.cpp:65-86)
            for (int s = o.nx * o.ny; s < n_surf_plane; s++) {
                o.cur[(size_t)s * o.G + g] = 0.0;
                o.sfl[(size_t)s * o.G + g] = 0.0;
            }
            sweep1g(o, g, true);
            o.has_radial = true;
        } else {
            sweep1g(o, g, false);
        }
    }
}

// MoCSweeper_2D3D::sweep (moc_sweeper_2d3d.cpp:48-99): the inner iterations
// of MoCSweeper::sweep, the last one with the CurrentCorrections worker after
// the pin-cell cross sections are homogenized from the current flux.
void mocsweep_2d3d(Oracle &o, int g)
{
    for (int f = 0; f < o.F; f++) o.xstr[f] = o.xs_tr[o.fsr_mat[f] * o.G + g];
    const int n_surf_plane = o.nx * o.ny + (o.nx + 1) * o.ny + (o.ny + 1) * o.nx;
    for (int inner = 0; inner < o.o.n_inner; inner++) {
        source_self_scatter(o, g);
        if (inner == o.o.n_inner - 1) {
            for (int s = o.nx * o.ny; s < n_surf_plane; s++) {
                o.cur[(size_t)s * o.G + g] = 0.0;
                o.sfl[(size_t)s * o.G + g] = 0.0;
            }
            homogenize(o);
            sweep1g(o, g, true, 2);
            o.has_radial = true;
        } else {
            sweep1g(o, g, false, 0);
        }
    }
}

// MoCSweeper::update_incoming_flux (moc_sweeper.cpp:255-291) through
// update_incoming_generic (moc_sweeper.hpp:198-233): the incoming flux at
// each ray's entry point scaled by the ratio of the CMFD partial currents of
// the latest solve to those of the previous one (or set from the latest
// ones after the first solve). The reference indexes the partial currents
// with the ray's first coarse cell (cm_cell_fw / cm_cell_bw), restated as is.
void update_incoming_flux(Oracle &o)
{
    const int G = o.G;
    auto part = [&](const std::vector<double> &p, int is, int g) {
        return 2.0 * (p[((size_t)is * G + g) * 2 + 0] + p[((size_t)is * G + g) * 2 + 1]);
    };
    auto update = [&](double in, int is, int g) {
        if (o.has_old_partial) {
            const double pn = part(o.part, is, g), po = part(o.part_old, is, g);
            return po > 0.0 ? in * (pn / po) : in;
        }
        return (2.0 / FPI) * (o.part[((size_t)is * G + g) * 2 + 0] + o.part[((size_t)is * G + g) * 2 + 1]);
    };
    for (int g = 0; g < G; g++)
        for (int iang = 0; iang < 2 * o.nd; iang++) {
            const int a = iang / o.NP;
            const int iang2 = reverse_ang(o, iang);
            double *bc_fw = &o.bc_in[(size_t)g * o.bcpg + o.off[iang * 3]];
            double *bc_bw = &o.bc_in[(size_t)g * o.bcpg + o.off[iang2 * 3]];
            for (int ir = o.azi_ray_off[a]; ir < o.azi_ray_off[a + 1]; ir++) {
                const int is1 = o.ray_cm_cell[2 * ir], is2 = o.ray_cm_cell[2 * ir + 1];
                const int bc1 = o.ray_bc[2 * ir], bc2 = o.ray_bc[2 * ir + 1];
                bc_fw[bc1] = update(bc_fw[bc1], is1, g);
                bc_bw[bc2] = update(bc_bw[bc2], is2, g);
            }
        }
}

// TransportSweeper::calc_fission_source (transport_sweeper.cpp:119-134)
void calc_fission_source(Oracle &o, double k, std::vector<double> &out)
{
    double rk = 1.0 / k;
    out.assign(o.F, 0.0);
    for (int f = 0; f < o.F; f++) {
        int m = o.fsr_mat[f];
        for (int g = 0; g < o.G; g++) out[f] += rk * o.xs_nf[m * o.G + g] * o.flux[(size_t)f * o.G + g];
    }
}

// TransportSweeper::total_fission (transport_sweeper.cpp:104-117)
double total_fission(const Oracle &o, bool old)
{
    const std::vector<double> &phi = old ? o.flux_old : o.flux;
    double t = 0.0;
    for (int m = 0; m < o.p.n_mat; m++)
        for (int g = 0; g < o.G; g++) {
            double nf = o.xs_nf[m * o.G + g];
            for (int f = 0; f < o.F; f++)
                if (o.fsr_mat[f] == m) t += phi[(size_t)f * o.G + g] * o.fsr_vol[f] * nf;
        }
    return t;
}

// FixedSourceSolver::step (fixed_source_solver.cpp:102-117)
void fss_step(Oracle &o, bool with_current)
{
    o.flux_old = o.flux;
    for (int g = 0; g < o.G; g++) {
        source_initialize(o);
        if (o.has_fs) source_fission(o, g);
        source_in_scatter(o, g, o.o.energy_mode == 1 ? o.flux_old : o.flux);
        mocsweep(o, g, with_current);
    }
}

// MoCSweeper::get_pin_flux_1g (PIN_PLANE), moc_sweeper.cpp:293-326
void get_pin_flux_1g(const Oracle &o, int g, double *pin)
{
    int npin = o.nx * o.ny;
    for (int i = 0; i < npin; i++) pin[i] = 0.0;
    for (int ip = 0; ip < npin; ip++) {
        double v = 0.0, pf = 0.0;
        for (int r = o.pin_fsr_off[ip]; r < o.pin_fsr_off[ip + 1]; r++) {
            v += o.fsr_vol[r];
            pf += o.flux[(size_t)r * o.G + g] * o.fsr_vol[r];
        }
        pf /= v;
        pin[o.pin_cell[ip]] += pf;
    }
}

// MoCSweeper::set_pin_flux_1g (PIN_PLANE), moc_sweeper.cpp:360-433
double set_pin_flux_1g(Oracle &o, int g, const double *pin)
{
    int npin = o.nx * o.ny;
    for (int i = 0; i < npin; i++)
        if (pin[i] <= 0.0) throw std::domain_error("Negative or zero input flux");
    double resid = 0.0;
    for (int ip = 0; ip < npin; ip++) {
        int ic = o.pin_cell[ip];
        double fm = 0.0;
        int r0 = o.pin_fsr_off[ip], r1 = o.pin_fsr_off[ip + 1];
        for (int r = r0; r < r1; r++) fm += o.flux[(size_t)r * o.G + g] * o.fsr_area[r];
        fm /= o.pin_area[ip];
        double e = pin[ic] - fm;
        double f = pin[ic] / fm;
        for (int r = r0; r < r1; r++) o.flux[(size_t)r * o.G + g] *= f;
        resid += e * e;
    }
    return std::sqrt(resid) / npin;
}

// ---- CMFD (src/core/cmfd.cpp, xs_mesh_homogenized.cpp) ---------------------
// XSMeshHomogenized::homogenize_region_flux (xs_mesh_homogenized.cpp:261-363)
void homogenize(Oracle &o)
{
    const int G = o.G, npin = o.nx * o.ny;
    o.h_tr.assign((size_t)npin * G, 0.0);
    o.h_nf.assign((size_t)npin * G, 0.0);
    o.h_f.assign((size_t)npin * G, 0.0);
    o.h_ch.assign((size_t)npin * G, 0.0);
    o.h_rm.assign((size_t)npin * G, 0.0);
    o.h_sc.assign((size_t)npin * G * G, 0.0);
    for (int ip = 0; ip < npin; ip++) {
        int cell = o.pin_cell[ip];
        int r0 = o.pin_fsr_off[ip], r1 = o.pin_fsr_off[ip + 1];
        int nr = r1 - r0;
        std::vector<double> fsv(nr, 0.0);
        for (int g = 0; g < G; g++)
            for (int r = r0; r < r1; r++) {
                int m = o.fsr_mat[r];
                fsv[r - r0] += o.xs_nf[m * G + g] * o.flux[(size_t)r * G + g] * o.fsr_area[r];
            }
        double fs_sum = 0.0;
        for (double v : fsv) fs_sum += v;
        std::vector<double> scat((size_t)G * G, 0.0);
        for (int g = 0; g < G; g++) {
            double fvs = 0.0, tr = 0.0, nf = 0.0, fi = 0.0, ch = 0.0;
            std::vector<double> scatsum(G, 0.0);
            for (int r = r0; r < r1; r++) {
                int m = o.fsr_mat[r];
                double v = o.fsr_area[r];
                double fl = o.flux[(size_t)r * G + g];
                fvs += v * fl;
                tr += v * fl * o.xs_tr[m * G + g];
                nf += v * fl * o.xs_nf[m * G + g];
                fi += v * fl * o.xs_f[m * G + g];
                ch += fsv[r - r0] * o.xs_ch[m * G + g];
                const double *row = &o.xs_sc[(size_t)m * G * G + (size_t)g * G];
                for (int gg = 0; gg < G; gg++) {
                    double fgg = o.flux[(size_t)r * G + gg];
                    scatsum[gg] += fgg * v;
                    scat[(size_t)g * G + gg] += row[gg] * v * fgg;
                }
            }
            for (int gg = 0; gg < G; gg++)
                if (scat[(size_t)g * G + gg] > 0.0) scat[(size_t)g * G + gg] /= scatsum[gg];
            o.h_tr[(size_t)cell * G + g] = tr / fvs;
            o.h_nf[(size_t)cell * G + g] = nf / fvs;
            o.h_f[(size_t)cell * G + g] = fi / fvs;
            o.h_ch[(size_t)cell * G + g] = (fs_sum > 0.0) ? ch / fs_sum : ch;
        }
        for (int g = 0; g < G; g++) {
            for (int gg = 0; gg < G; gg++) o.h_sc[(size_t)cell * G * G + (size_t)g * G + gg] = scat[(size_t)g * G + gg];
            o.h_rm[(size_t)cell * G + g] = o.h_tr[(size_t)cell * G + g] - scat[(size_t)g * G + g];
        }
    }
}

double cell_thickness(const Oracle &o, int cell, int normal)
{
    if (normal == 0) return dx_of(o, cell % o.nx);
    if (normal == 1) return dy_of(o, cell / o.nx);
    return o.p.hz;
}

// CMFD::setup_solve (cmfd.cpp:333-553)
void cmfd_setup(Oracle &o)
{
    const int G = o.G, nc = o.n_cell;
    // Mesh::boundary_array: [normal][0 = low side, 1 = high side]
    int bca[3][2] = {{o.p.bc[WEST], o.p.bc[EAST]}, {o.p.bc[SOUTH], o.p.bc[NORTH]}, {o.p.bc[BOTTOM], o.p.bc[TOP]}};
    o.d_tilde.assign((size_t)o.n_surf * G, 0.0);
    o.d_hat.assign((size_t)o.n_surf * G, 0.0);
    o.s_tilde.assign((size_t)o.n_surf * G, 0.0);
    o.s_hat.assign((size_t)o.n_surf * G, 0.0);
    o.m_rowptr.assign(G, {});
    o.m_col.assign(G, {});
    o.m_val.assign(G, {});
    for (int g = 0; g < G; g++) {
        std::vector<double> dco(nc), xsrm(nc);
        for (int c = 0; c < nc; c++) {
            dco[c] = 1.0 / (3.0 * o.h_tr[(size_t)c * G + g]);
            xsrm[c] = o.h_rm[(size_t)c * G + g];
        }
        for (int is = 0; is < o.n_surf; is++) {
            auto cells = coarse_neigh_cells(o, is);
            int norm = surface_normal(o, is);
            double d1 = 0.0, d2 = 0.0;
            auto bdiff = [](int kind) {
                if (kind == BC_REFLECT) return 0.0 / 2.0;
                if (kind == BC_VACUUM) return 0.5 / 2.0;
                throw std::runtime_error("Unsupported boundary type");
            };
            d1 = cells.first > -1 ? dco[cells.first] / cell_thickness(o, cells.first, norm) : bdiff(bca[norm][0]);
            d2 = cells.second > -1 ? dco[cells.second] / cell_thickness(o, cells.second, norm) : bdiff(bca[norm][1]);
            double dt = 2.0 * d1 * d2 / (d1 + d2);
            double st = (d1 > 0.0) ? d1 / (d1 + d2) : d2 / (d1 + d2);
            o.d_tilde[(size_t)is * G + g] = dt;
            o.s_tilde[(size_t)is * G + g] = st;
            bool have = (norm == 2) ? false : o.has_radial;
            if (have) {
                double j = o.cur[(size_t)is * G + g];
                double sf = o.sfl[(size_t)is * G + g];
                double fl = cells.first >= 0 ? o.cflux[(size_t)cells.first * G + g] : 0.0;
                double fr = cells.second >= 0 ? o.cflux[(size_t)cells.second * G + g] : 0.0;
                double dh = (j + dt * (fr - fl)) / (fl + fr);
                if (!std::isfinite(dh)) dh = 0.0;
                o.d_hat[(size_t)is * G + g] = dh;
                o.s_hat[(size_t)is * G + g] = (cells.first >= 0) ? (sf - st * fl - (1.0 - st) * fr) / (fl + fr)
                                                                 : (sf - st * fr) / fr;
            }
        }
        // matrix: rows in cell order, columns sorted (diag + x/y neighbours)
        auto &rp = o.m_rowptr[g];
        auto &cl = o.m_col[g];
        auto &vl = o.m_val[g];
        rp.push_back(0);
        static const int order[6] = {EAST, NORTH, WEST, SOUTH, TOP, BOTTOM};
        for (int i = 0; i < nc; i++) {
            std::vector<std::pair<int, double>> row;
            double v = dx_of(o, i % o.nx) * dy_of(o, i / o.nx) * o.p.hz * xsrm[i];
            for (int s : order) {
                int surf = coarse_surf(o, i, s);
                double a = coarse_area_cell(o, i, s);
                double dh = o.d_hat[(size_t)surf * G + g];
                if (s == WEST || s == SOUTH || s == BOTTOM) dh = -dh;
                v += a * (o.d_tilde[(size_t)surf * G + g] + dh);
            }
            row.push_back({i, v});
            for (int s : order) {
                int n = coarse_neighbor(o, i, s);
                if (n < 0) continue;
                int surf = coarse_surf(o, i, s);
                double a = coarse_area_cell(o, i, s);
                double dh = o.d_hat[(size_t)surf * G + g];
                if (s == WEST || s == SOUTH || s == BOTTOM) dh = -dh;
                row.push_back({n, a * (dh - o.d_tilde[(size_t)surf * G + g])});
            }
            std::sort(row.begin(), row.end());
            for (auto &e : row) {
                cl.push_back(e.first);
                vl.push_back(e.second);
            }
            rp.push_back((int)cl.size());
        }
    }
}

void spmv(const std::vector<int> &rp, const std::vector<int> &cl, const std::vector<double> &vl,
          const double *x, double *y, int n)
{
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += vl[k] * x[cl[k]];
        y[i] = s;
    }
}

// Right-preconditioned BiCGSTAB with a Jacobi (diagonal) preconditioner and
// the relative stopping test |r| <= tol |b|, restarting when rho becomes
// numerically zero: the algorithm of van der Vorst (1992) as Eigen implements
// it for BiCGSTAB<SparseMatrix<double>> (cmfd.hpp:221, cmfd.cpp:291).
int bicgstab(int n, const std::vector<int> &rp, const std::vector<int> &cl, const std::vector<double> &vl,
             const double *b, double *x, double tol, int max_iter, double *err)
{
    std::vector<double> invdiag(n, 1.0);
    for (int i = 0; i < n; i++)
        for (int k = rp[i]; k < rp[i + 1]; k++)
            if (cl[k] == i && vl[k] != 0.0) invdiag[i] = 1.0 / vl[k];
    auto dot = [n](const std::vector<double> &a, const std::vector<double> &c) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += a[i] * c[i];
        return s;
    };
    std::vector<double> r(n), r0(n), v(n, 0.0), p(n, 0.0), y(n), z(n), s(n), t(n), tmp(n);
    spmv(rp, cl, vl, x, tmp.data(), n);
    for (int i = 0; i < n; i++) r[i] = b[i] - tmp[i];
    r0 = r;
    double r0_sq = dot(r0, r0);
    double rhs_sq = 0.0;
    for (int i = 0; i < n; i++) rhs_sq += b[i] * b[i];
    if (rhs_sq == 0.0) {
        for (int i = 0; i < n; i++) x[i] = 0.0;
        if (err) *err = 0.0;
        return 0;
    }
    double rho = 1, alpha = 1, w = 1;
    double tol2 = tol * tol * rhs_sq;
    double eps2 = std::numeric_limits<double>::epsilon() * std::numeric_limits<double>::epsilon();
    int i = 0, restarts = 0;
    while (dot(r, r) > tol2 && i < max_iter) {
        double rho_old = rho;
        rho = dot(r0, r);
        if (std::abs(rho) < eps2 * r0_sq) {
            spmv(rp, cl, vl, x, tmp.data(), n);
            for (int k = 0; k < n; k++) r[k] = b[k] - tmp[k];
            r0 = r;
            rho = r0_sq = dot(r, r);
            if (restarts++ == 0) i = 0;
        }
        double beta = (rho / rho_old) * (alpha / w);
        for (int k = 0; k < n; k++) p[k] = r[k] + beta * (p[k] - w * v[k]);
        for (int k = 0; k < n; k++) y[k] = invdiag[k] * p[k];
        spmv(rp, cl, vl, y.data(), v.data(), n);
        alpha = rho / dot(r0, v);
        for (int k = 0; k < n; k++) s[k] = r[k] - alpha * v[k];
        for (int k = 0; k < n; k++) z[k] = invdiag[k] * s[k];
        spmv(rp, cl, vl, z.data(), t.data(), n);
        double tt = dot(t, t);
        w = (tt > 0.0) ? dot(t, s) / tt : 0.0;
        for (int k = 0; k < n; k++) x[k] += alpha * y[k] + w * z[k];
        for (int k = 0; k < n; k++) r[k] = s[k] - w * t[k];
        ++i;
    }
    if (err) *err = std::sqrt(dot(r, r) / rhs_sq);
    return i;
}

struct CmfdSource {
    std::vector<double> b;
};

// CMFD source for group g: V * (chi fs + in-scatter), cmfd.cpp:225-231
void cmfd_source(const Oracle &o, int g, const std::vector<double> &fsc, std::vector<double> &b)
{
    const int G = o.G;
    b.assign(o.n_cell, 0.0);
    for (int c = 0; c < o.n_cell; c++) b[c] += o.h_ch[(size_t)c * G + g] * fsc[c];
    for (int c = 0; c < o.n_cell; c++) {
        const double *row = &o.h_sc[(size_t)c * G * G + (size_t)g * G];
        for (int gg = 0; gg < G; gg++) {
            if (gg == g) continue;
            b[c] += row[gg] * o.cflux[(size_t)c * G + gg];
        }
    }
    for (int c = 0; c < o.n_cell; c++) b[c] *= dx_of(o, c % o.nx) * dy_of(o, c / o.nx) * o.p.hz;
}

double cmfd_residual_g(const Oracle &o, int g, const std::vector<double> &b)
{
    std::vector<double> x(o.n_cell), ax(o.n_cell);
    for (int c = 0; c < o.n_cell; c++) x[c] = o.cflux[(size_t)c * o.G + g];
    spmv(o.m_rowptr[g], o.m_col[g], o.m_val[g], x.data(), ax.data(), o.n_cell);
    double s = 0.0;
    for (int c = 0; c < o.n_cell; c++) s += (ax[c] - b[c]) * (ax[c] - b[c]);
    return s;
}

void cmfd_fission_source(const Oracle &o, double k, std::vector<double> &fsc)
{
    fsc.assign(o.n_cell, 0.0);
    for (int c = 0; c < o.n_cell; c++)
        for (int g = 0; g < o.G; g++) fsc[c] += o.h_nf[(size_t)c * o.G + g] * o.cflux[(size_t)c * o.G + g];
    double rk = 1.0 / k;
    for (auto &v : fsc) v *= rk;
}

double cmfd_total_fission(const Oracle &o)
{
    double f = 0.0;
    for (int c = 0; c < o.n_cell; c++)
        for (int g = 0; g < o.G; g++) f += o.h_nf[(size_t)c * o.G + g] * o.cflux[(size_t)c * o.G + g];
    return f;
}

// CMFD::store_currents (cmfd.cpp:555-638)
void cmfd_store_currents(Oracle &o)
{
    const int G = o.G;
    for (int g = 0; g < G; g++) {
        for (int is = 0; is < o.n_surf; is++) {
            o.part_old[((size_t)is * G + g) * 2 + 0] = o.part[((size_t)is * G + g) * 2 + 0];
            o.part_old[((size_t)is * G + g) * 2 + 1] = o.part[((size_t)is * G + g) * 2 + 1];
        }
        o.has_old_partial = o.n_solve > 0;
        for (int is = 0; is < o.n_surf; is++) {
            auto cells = coarse_neigh_cells(o, is);
            double fr = cells.second >= 0 ? o.cflux[(size_t)cells.second * G + g] : 0.0;
            double fl = cells.first >= 0 ? o.cflux[(size_t)cells.first * G + g] : 0.0;
            double dh = o.d_hat[(size_t)is * G + g], dt = o.d_tilde[(size_t)is * G + g];
            double j = -dt * (fr - fl) + dh * (fr + fl);
            double sh = o.s_hat[(size_t)is * G + g], st = o.s_tilde[(size_t)is * G + g];
            double sf = (cells.first >= 0) ? st * fl + (1.0 - st) * fr + sh * (fl + fr) : st * fr + sh * (fl + fr);
            o.sfl[(size_t)is * G + g] = sf;
            o.part[((size_t)is * G + g) * 2 + 0] = 0.25 * sf + 0.5 * j;
            o.part[((size_t)is * G + g) * 2 + 1] = 0.25 * sf - 0.5 * j;
            o.cur[(size_t)is * G + g] = j;
        }
    }
}

// CMFD::solve (cmfd.cpp:171-279)
int cmfd_solve(Oracle &o, double &k)
{
    const int G = o.G, nc = o.n_cell;
    // <cmfd negative_fixup>: no negative coarse flux into the solve (cmfd.cpp:175-184)
    if (o.o.cmfd_negative_fixup)
        for (auto &v : o.cflux) v = std::max(v, 0.0);
    homogenize(o);
    cmfd_setup(o);
    double k_old = k;
    std::vector<double> fsc, fs_old, b;
    cmfd_fission_source(o, k, fsc);
    double tfis = cmfd_total_fission(o);
    double r0 = 0.0;
    for (int g = 0; g < G; g++) {
        cmfd_source(o, g, fsc, b);
        r0 += cmfd_residual_g(o, g, b);
    }
    r0 = std::sqrt(r0) / (G * nc);
    double tol = o.o.cmfd_resid_reduction * r0;
    int iter = 0;
    double psi_err = 1.0, ri = 0.0;
    while (true) {
        iter++;
        fs_old = fsc;
        cmfd_fission_source(o, k, fsc);
        double tfis_old = tfis;
        ri = 0.0;
        for (int g = 0; g < G; g++) {
            cmfd_source(o, g, fsc, b);
            ri += cmfd_residual_g(o, g, b);
            std::vector<double> x(nc);
            for (int c = 0; c < nc; c++) x[c] = o.cflux[(size_t)c * G + g];
            bicgstab(nc, o.m_rowptr[g], o.m_col[g], o.m_val[g], b.data(), x.data(), tol, 150, nullptr);
            for (int c = 0; c < nc; c++) o.cflux[(size_t)c * G + g] = x[c];
        }
        ri = std::sqrt(ri) / (nc * G);
        tfis = cmfd_total_fission(o);
        k_old = k;
        k = k * tfis / tfis_old;
        psi_err = 0.0;
        for (int c = 0; c < nc; c++) psi_err += (fsc[c] - fs_old[c]) * (fsc[c] - fs_old[c]);
        psi_err = std::sqrt(psi_err);
        if (((std::abs(k - k_old) < o.o.cmfd_k_tol) && (psi_err < o.o.cmfd_psi_tol) &&
             (ri / r0 < o.o.cmfd_resid_reduction)) ||
            (iter > o.o.cmfd_max_iter))
            break;
    }
    for (auto &v : o.cflux)
        if (v < 0.0) v = -v;
    cmfd_store_currents(o);
    o.n_solve++;
    return iter;
}

// EigenSolver::do_cmfd (eigen_solver.cpp:263-291)
int do_cmfd(Oracle &o, double &k)
{
    std::vector<double> pin(o.n_cell);
    for (int g = 0; g < o.G; g++) {
        get_pin_flux_1g(o, g, pin.data());
        for (int c = 0; c < o.n_cell; c++) o.cflux[(size_t)c * o.G + g] = pin[c];
    }
    int it = cmfd_solve(o, k);
    for (int g = 0; g < o.G; g++) {
        for (int c = 0; c < o.n_cell; c++) pin[c] = o.cflux[(size_t)c * o.G + g];
        set_pin_flux_1g(o, g, pin.data());
    }
    return it;
}

// util/utils.hpp:28-51 Normalize: scale to sum = number of positive entries
void normalize(std::vector<double> &v)
{
    int n = 0;
    double sum = 0.0;
    for (double x : v) {
        if (x > 0.0) n++;
        sum += x;
    }
    double f = (double)n / sum;
    for (double &x : v) x *= f;
}

struct EigenState {
    double k = 1.0, k_prev = 1.0, err_k = 0.0, err_psi = 0.0;
    int iterations = 0, converged = 0;
    int n_fissile = 0;
    std::vector<double> fs_prev;
};

// EigenSolver::step + the convergence bookkeeping of EigenSolver::solve
// (eigen_solver.cpp:225-253, 162-217)
void eigen_iterate(Oracle &o, EigenState &st)
{
    if (o.o.use_cmfd) {
        int it = do_cmfd(o, st.k);
        o.last_cmfd_iters = it;
    }
    calc_fission_source(o, st.k, o.fs);
    o.has_fs = true;
    st.fs_prev = o.fs;
    fss_step(o, o.o.use_cmfd != 0);
    double tf1 = total_fission(o, false);
    double tf2 = total_fission(o, true);
    st.k_prev = st.k;
    st.k = st.k * tf1 / tf2;
    calc_fission_source(o, st.k, o.fs);
    st.err_k = std::fabs(st.k - st.k_prev);
    std::vector<double> a = o.fs, b = st.fs_prev;
    normalize(a);
    normalize(b);
    double e = 0.0;
    for (int i = 0; i < o.F; i++) e += (a[i] - b[i]) * (a[i] - b[i]);
    st.err_psi = std::sqrt(e / st.n_fissile);
    st.iterations++;
}

int count_fissile(const Oracle &o)
{
    int n = 0;
    for (int f = 0; f < o.F; f++) {
        int m = o.fsr_mat[f];
        for (int g = 0; g < o.G; g++)
            if (o.xs_nf[m * o.G + g] > 0.0) {
                n++;
                break;
            }
    }
    return n;
}

}  // namespace

struct oracle_handle {
    Oracle o;
    EigenState es;
};

extern "C" {

void oracle_default_options(oracle_options *o)
{
    *o = oracle_options{};
    o->n_inner = 1;
    o->gs_boundary = 1;
    o->energy_mode = 0;
    o->exact_exp = 0;
    o->use_cmfd = 0;
    o->cmfd_max_iter = 100;
    o->cmfd_k_tol = 1e-6;
    o->cmfd_psi_tol = 1e-5;
    o->cmfd_resid_reduction = 1e-3;
    o->n_threads = 0;
    o->cmfd_negative_fixup = 0;
}

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_create(const mocgpu_problem *pp, const oracle_options *opt, oracle_handle **out)
{
    try {
        auto h = new oracle_handle();
        Oracle &o = h->o;
        o.p = *pp;
        o.o = *opt;
        const mocgpu_problem &p = *pp;
        o.F = p.n_fsr;
        o.G = p.n_group;
        o.A = p.n_azi;
        o.NP = p.n_polar;
        o.nd = p.n_azi * p.n_polar / 2;
        int na = 2 * o.nd;
        copy_vec(o.ang_weight, p.ang_weight, na);
        copy_vec(o.ang_sin, p.ang_sin_theta, na);
        copy_vec(o.ang_rsin, p.ang_rsin_theta, na);
        copy_vec(o.ang_ox, p.ang_ox, na);
        copy_vec(o.ang_oy, p.ang_oy, na);
        copy_vec(o.ang_alpha, p.ang_alpha, na);
        copy_vec(o.azi_ray_off, p.azi_ray_off, o.A + 1);
        copy_vec(o.azi_nx, p.azi_nx, o.A);
        copy_vec(o.azi_ny, p.azi_ny, o.A);
        copy_vec(o.azi_spacing, p.azi_spacing, o.A);
        copy_vec(o.ray_seg_off, p.ray_seg_off, p.n_ray + 1);
        copy_vec(o.ray_bc, p.ray_bc, 2 * (size_t)p.n_ray);
        copy_vec(o.seg_fsr, p.seg_fsr, p.n_seg);
        copy_vec(o.seg_len, p.seg_len, p.n_seg);
        copy_vec(o.ray_cm_cell, p.ray_cm_cell, 2 * (size_t)p.n_ray);
        copy_vec(o.ray_cm_surf, p.ray_cm_surf, 2 * (size_t)p.n_ray);
        copy_vec(o.ray_cm_off, p.ray_cm_off, p.n_ray + 1);
        copy_vec(o.cm_data, p.cm_data, 4 * (size_t)o.ray_cm_off.back());
        copy_vec(o.fsr_vol, p.fsr_vol, o.F);
        copy_vec(o.fsr_area, p.fsr_area, o.F);
        copy_vec(o.fsr_mat, p.fsr_mat, o.F);
        copy_vec(o.xs_tr, p.xs_tr, (size_t)p.n_mat * o.G);
        copy_vec(o.xs_nf, p.xs_nf, (size_t)p.n_mat * o.G);
        copy_vec(o.xs_f, p.xs_f, (size_t)p.n_mat * o.G);
        copy_vec(o.xs_ch, p.xs_ch, (size_t)p.n_mat * o.G);
        copy_vec(o.xs_sc, p.xs_sc, (size_t)p.n_mat * o.G * o.G);
        o.nx = p.nx;
        o.ny = p.ny;
        copy_vec(o.x_vec, p.x_vec, o.nx + 1);
        copy_vec(o.y_vec, p.y_vec, o.ny + 1);
        copy_vec(o.pin_fsr_off, p.pin_fsr_off, o.nx * o.ny + 1);
        copy_vec(o.pin_cell, p.pin_cell, o.nx * o.ny);
        copy_vec(o.pin_area, p.pin_area, o.nx * o.ny);
        // re-point the view at the copies
        o.p.seg_fsr = o.seg_fsr.data();
        for (int r = 0; r < p.n_ray; r++)
            o.max_seg = std::max(o.max_seg, (int)(o.ray_seg_off[r + 1] - o.ray_seg_off[r]));
        // boundary sizes (moc_sweeper.cpp:38-52): per angle {Ny, Nx, 0}
        o.fsize.assign(4 * o.nd * 3, 0);
        for (int iang = 0; iang < na; iang++) {
            int a = iang / o.NP;
            int i2 = reverse_ang(o, iang);
            o.fsize[iang * 3 + 0] = o.fsize[i2 * 3 + 0] = o.azi_ny[a];
            o.fsize[iang * 3 + 1] = o.fsize[i2 * 3 + 1] = o.azi_nx[a];
        }
        o.off.assign(4 * o.nd * 3, 0);
        int acc = 0;
        for (int iang = 0; iang < 4 * o.nd; iang++)
            for (int f = 0; f < 3; f++) {
                o.off[iang * 3 + f] = acc;
                acc += o.fsize[iang * 3 + f];
            }
        o.bcpg = acc;
        o.bc_in.assign((size_t)o.G * o.bcpg, 0.0);
        o.bc_out.assign(o.bcpg, 0.0);
        o.flux.assign((size_t)o.F * o.G, 1.0);
        o.flux_old = o.flux;
        o.xstr.assign(o.F, 0.0);
        o.source_1g.assign(o.F, 0.0);
        o.q.assign(o.F, 0.0);
        o.fs.assign(o.F, 0.0);
        o.n_cell = o.nx * o.ny;
        o.n_surf = (o.nx + 1) * o.ny + (o.ny + 1) * o.nx + 2 * o.nx * o.ny;
        o.cur.assign((size_t)o.n_surf * o.G, 0.0);
        o.sfl.assign((size_t)o.n_surf * o.G, 0.0);
        o.part.assign((size_t)o.n_surf * o.G * 2, 0.0);
        o.part_old = o.part;
        o.alpha.assign((size_t)o.G * 4 * o.nd * o.nx * o.ny * 2, 0.5);
        o.beta.assign((size_t)o.G * 4 * o.nd * o.nx * o.ny, 1.0);
        o.corr_resid.assign((size_t)o.G * 3, 0.0);
        o.cflux.assign((size_t)o.n_cell * o.G, 1.0);
        h->es.n_fissile = count_fissile(o);
        *out = h;
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MOCGPU_ERR_ARG;
    }
}

void oracle_destroy(oracle_handle *h) { delete h; }

int oracle_initialize(oracle_handle *h)
{
    Oracle &o = h->o;
    std::fill(o.cflux.begin(), o.cflux.end(), 1.0);
    std::fill(o.flux.begin(), o.flux.end(), 1.0);
    std::fill(o.flux_old.begin(), o.flux_old.end(), 1.0);
    bc_initialize_scalar(o, 1.0 / FPI);
    return 0;
}

int oracle_set_flux(oracle_handle *h, const double *phi)
{
    std::copy(phi, phi + h->o.flux.size(), h->o.flux.begin());
    return 0;
}
int oracle_get_flux(oracle_handle *h, double *phi)
{
    std::copy(h->o.flux.begin(), h->o.flux.end(), phi);
    return 0;
}
int64_t oracle_boundary_size(const oracle_handle *h) { return (int64_t)h->o.bc_in.size(); }
int oracle_get_boundary(oracle_handle *h, double *bc)
{
    std::copy(h->o.bc_in.begin(), h->o.bc_in.end(), bc);
    return 0;
}
int oracle_set_boundary(oracle_handle *h, const double *bc)
{
    std::copy(bc, bc + h->o.bc_in.size(), h->o.bc_in.begin());
    return 0;
}
int oracle_store_old_flux(oracle_handle *h)
{
    h->o.flux_old = h->o.flux;
    return 0;
}
int oracle_calc_fission_source(oracle_handle *h, double k, double *fs_out)
{
    calc_fission_source(h->o, k, h->o.fs);
    h->o.has_fs = true;
    if (fs_out) std::copy(h->o.fs.begin(), h->o.fs.end(), fs_out);
    return 0;
}
int oracle_set_fission_source(oracle_handle *h, const double *fs)
{
    std::copy(fs, fs + h->o.F, h->o.fs.begin());
    h->o.has_fs = true;
    return 0;
}
int oracle_build_source(oracle_handle *h, int32_t g)
{
    Oracle &o = h->o;
    source_initialize(o);
    if (o.has_fs) source_fission(o, g);
    source_in_scatter(o, g, o.o.energy_mode == 1 ? o.flux_old : o.flux);
    return 0;
}
int oracle_get_source(oracle_handle *h, double *src_1g, double *qbar)
{
    if (src_1g) std::copy(h->o.source_1g.begin(), h->o.source_1g.end(), src_1g);
    if (qbar) std::copy(h->o.q.begin(), h->o.q.end(), qbar);
    return 0;
}
int oracle_sweep(oracle_handle *h, int32_t g, int32_t with_current)
{
    try {
        mocsweep(h->o, g, with_current != 0);
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MOCGPU_ERR_ARG;
    }
}
int oracle_fss_step(oracle_handle *h, int32_t with_current)
{
    try {
        fss_step(h->o, with_current != 0);
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MOCGPU_ERR_ARG;
    }
}
int oracle_total_fission(oracle_handle *h, int32_t old, double *out)
{
    *out = total_fission(h->o, old != 0);
    return 0;
}
int oracle_flux_residual(oracle_handle *h, double *out)
{
    double r = 0.0;
    for (size_t i = 0; i < h->o.flux.size(); i++) {
        double e = h->o.flux[i] - h->o.flux_old[i];
        r += e * e;
    }
    *out = std::sqrt(r);
    return 0;
}
int32_t oracle_n_surf(const oracle_handle *h) { return h->o.n_surf; }
int oracle_get_currents(oracle_handle *h, int32_t g, double *cur, double *sfl)
{
    Oracle &o = h->o;
    for (int s = 0; s < o.n_surf; s++) {
        if (cur) cur[s] = o.cur[(size_t)s * o.G + g];
        if (sfl) sfl[s] = o.sfl[(size_t)s * o.G + g];
    }
    return 0;
}
int oracle_get_pin_flux_1g(oracle_handle *h, int32_t g, double *pin)
{
    get_pin_flux_1g(h->o, g, pin);
    return 0;
}
int oracle_set_pin_flux_1g(oracle_handle *h, int32_t g, const double *pin, double *resid)
{
    try {
        double r = set_pin_flux_1g(h->o, g, pin);
        if (resid) *resid = r;
        return 0;
    } catch (const std::domain_error &e) {
        g_err = e.what();
        return MOCGPU_ERR_NONPOS_FLUX;
    }
}
int oracle_cmfd_step(oracle_handle *h, double *k, int32_t *iters)
{
    try {
        int it = do_cmfd(h->o, *k);
        if (iters) *iters = it;
        return 0;
    } catch (const std::domain_error &e) {
        g_err = e.what();
        return MOCGPU_ERR_NONPOS_FLUX;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MOCGPU_ERR_ARG;
    }
}
int oracle_cmfd_flux(oracle_handle *h, double *out)
{
    std::copy(h->o.cflux.begin(), h->o.cflux.end(), out);
    return (int)h->o.cflux.size();
}
int oracle_cmfd_matrix(oracle_handle *h, int32_t g, int32_t *row_ptr, int32_t *col, double *val, int32_t cap)
{
    Oracle &o = h->o;
    if (o.m_rowptr.empty()) return MOCGPU_ERR_ARG;
    const auto &rp = o.m_rowptr[g];
    if ((int)o.m_col[g].size() > cap) return MOCGPU_ERR_CAPACITY;
    std::copy(rp.begin(), rp.end(), row_ptr);
    std::copy(o.m_col[g].begin(), o.m_col[g].end(), col);
    std::copy(o.m_val[g].begin(), o.m_val[g].end(), val);
    return (int)o.m_col[g].size();
}

int oracle_eigen_iterate(oracle_handle *h, mocgpu_eigen_state *st)
{
    try {
        EigenState &es = h->es;
        es.k = st->k;
        eigen_iterate(h->o, es);
        st->k = es.k;
        st->k_prev = es.k_prev;
        st->err_k = es.err_k;
        st->err_psi = es.err_psi;
        st->iterations = es.iterations;
        if (st->k != st->k) {
            g_err = "Eigenvalue is not a number. Giving up.";
            return MOCGPU_ERR_NAN;
        }
        return 0;
    } catch (const std::domain_error &e) {
        g_err = e.what();
        return MOCGPU_ERR_NONPOS_FLUX;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MOCGPU_ERR_ARG;
    }
}

// EigenSolver::solve (eigen_solver.cpp:131-223)
int oracle_eigen_solve(oracle_handle *h, double k_tol, double psi_tol, int32_t max_iter, int32_t min_iter,
                       mocgpu_eigen_state *st, double *k_hist, int32_t hist_cap)
{
    Oracle &o = h->o;
    oracle_initialize(h);
    h->es = EigenState();
    h->es.n_fissile = count_fissile(o);
    mocgpu_eigen_state s{};
    s.k = 1.0;
    s.k_prev = 1.0;
    calc_fission_source(o, s.k, o.fs);
    o.has_fs = true;
    for (int n = 0; n < max_iter; n++) {
        int rc = oracle_eigen_iterate(h, &s);
        if (rc) {
            *st = s;
            return rc;
        }
        if (k_hist && n < hist_cap) k_hist[n] = s.k;
        if ((s.err_k < k_tol) && (s.err_psi < psi_tol) && (n >= min_iter)) {
            s.converged = 1;
            break;
        }
    }
    *st = s;
    return s.converged ? 0 : MOCGPU_ERR_MAXITER;
}

double oracle_exp_table(double v, int32_t n, double xmin, double xmax)
{
    ExpTable t(n, xmin, xmax);
    return t.exp(v);
}
double oracle_exp_table_node(int32_t i, int32_t n, double xmin, double xmax)
{
    ExpTable t(n, xmin, xmax);
    return t.d[i];
}

int oracle_time_sweeps(oracle_handle *h, int32_t n, double *ms_per_sweep)
{
    Oracle &o = h->o;
    std::vector<double> times;
    for (int it = 0; it < n + 1; it++) {
        auto t0 = std::chrono::steady_clock::now();
        for (int g = 0; g < o.G; g++) {
            for (int f = 0; f < o.F; f++) o.xstr[f] = o.xs_tr[o.fsr_mat[f] * o.G + g];
            sweep1g(o, g, false);
        }
        auto t1 = std::chrono::steady_clock::now();
        if (it > 0) times.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    std::sort(times.begin(), times.end());
    *ms_per_sweep = times.empty() ? 0.0 : times[times.size() / 2];
    return 0;
}

int oracle_bicgstab(int32_t n, const int32_t *row_ptr, const int32_t *col, const double *val, const double *b,
                    double *x, double tol, int32_t max_iter, int32_t *iters, double *err)
{
    std::vector<int> rp(row_ptr, row_ptr + n + 1), cl(col, col + row_ptr[n]);
    std::vector<double> vl(val, val + row_ptr[n]);
    int it = bicgstab(n, rp, cl, vl, b, x, tol, max_iter, err);
    if (iters) *iters = it;
    return 0;
}

int oracle_sweep_2d3d(oracle_handle *h, int32_t g)
{
    if (g < 0 || g >= h->o.G) return -1;
    mocsweep_2d3d(h->o, g);
    return 0;
}

int oracle_get_corrections(oracle_handle *h, double *alpha, double *beta)
{
    if (alpha) std::copy(h->o.alpha.begin(), h->o.alpha.end(), alpha);
    if (beta) std::copy(h->o.beta.begin(), h->o.beta.end(), beta);
    return 0;
}

int oracle_correction_residual(oracle_handle *h, int32_t g, double *out)
{
    if (g < 0 || g >= h->o.G) return -1;
    for (int k = 0; k < 3; k++) out[k] = h->o.corr_resid[(size_t)g * 3 + k];
    return 0;
}

int oracle_update_incoming_flux(oracle_handle *h)
{
    update_incoming_flux(h->o);
    return 0;
}

}  // extern "C"
