// Pin-cell CMFD in FP64 on the device.
//
// Restates mocc::CMFD (src/core/cmfd.cpp:171-638) with the homogenization of
// XSMeshHomogenized::homogenize_region_flux (xs_mesh_homogenized.cpp:261-363)
// and EigenSolver::do_cmfd (eigen_solver.cpp:263-291):
//
//   * one thread per (pin, group) homogenizes the fine flux and cross
//     sections, one per (surface, group) forms D-tilde / D-hat, one per
//     (cell, group) assembles the 5-point row;
//   * the coarse power iteration -- fission source, Gauss-Seidel sweep over
//     groups, one BiCGSTAB solve per group, the k update and the convergence
//     test -- runs in ONE workgroup: up to 512 lanes for meshes of up to 1024
//     cells, and the wide solve (512 or 1024 lanes) for larger meshes, for a
//     single problem and for each problem of a batch. Each lane owns K cells
//     (c = lane + j * lanes) and keeps their BiCGSTAB vectors in registers; a
//     matrix-vector product stages the operand in LDS, and dot products are
//     wave butterflies plus a fixed-order sum over the waves, so every lane
//     sees the same scalars and the iteration is deterministic. No host round
//     trip happens inside a CMFD solve.
#include "cmfd.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace mocc_amd {
namespace gpu {

namespace {

constexpr int kCmfdThreads = 512;
constexpr int kCmfdLdsVectors = 6;    // cell vectors of the solve in LDS (see cmfd_solve)

// Cells per lane of a mesh of nc cells on kCmfdThreads lanes: 1 or 2 for the
// meshes of up to 1024 cells this solve takes (larger ones run the wide
// solve); each LDS vector holds K * lanes entries, the cells and the dummies.
inline int cmfd_cells_per_lane(int nc) { return std::max(1, (nc + kCmfdThreads - 1) / kCmfdThreads); }
// Lanes of the solve: meshes of one cell per lane use the fewest whole waves
// that cover them (a barrier or a reduction then waits for no idle wave);
// larger meshes use kCmfdThreads. The lane of cell c is the same either way.
inline int cmfd_threads(int nc)
{
    return nc <= kCmfdThreads ? std::max(64, (nc + 63) / 64 * 64) : kCmfdThreads;
}
// LDS of the solve of up to 1024 cells: kCmfdLdsVectors vectors of K * lanes.
inline size_t cmfd_lds_bytes(int nc)
{
    return (size_t)kCmfdLdsVectors * cmfd_cells_per_lane(nc) * cmfd_threads(nc) * sizeof(double);
}
// Meshes of more than 1024 cells run the wide solve (k_cmfd_solve_wide): on
// 512 lanes with the rows in registers up to 3072 cells (3 to 6 cells per
// lane), on 1024 lanes re-reading the rows from cache beyond (4 to 6). Its LDS
// holds the two operand banks with a pad of nx entries at each end; its
// global scratch holds two vectors of K * lanes.
constexpr int kWideCells = 6144;
static_assert(kWideCells == CmfdSolver::kMaxCells, "the wide solve covers every single-problem mesh");
inline bool cmfd_wide(int nc) { return nc > 1024; }
inline int cmfd_wide_lanes(int nc) { return nc <= 3072 ? 512 : 1024; }
inline int cmfd_wide_k(int nc) { return (nc + cmfd_wide_lanes(nc) - 1) / cmfd_wide_lanes(nc); }
inline size_t cmfd_wide_lds_bytes(int nc, int nx)
{
    return (2 * (size_t)cmfd_wide_k(nc) * cmfd_wide_lanes(nc) + 2 * (size_t)nx) * sizeof(double);
}
enum { EAST = 0, NORTH = 1, WEST = 2, SOUTH = 3, TOP = 4, BOTTOM = 5 };

#define CMFD_CHECK(expr)                                                                          \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// --- coarse mesh numbering (Mesh::coarse_surf / coarse_neighbor /
//     coarse_neigh_cells / coarse_area, src/core/mesh.cpp:99-147, mesh.hpp) --
struct Grid {
    int nx, ny;
    double hz;
    const double *dx, *dy;
};

__device__ __forceinline__ int surf_normal(const Grid &m, int surf)
{
    const int nsp = m.nx * m.ny + (m.nx + 1) * m.ny + (m.ny + 1) * m.nx;
    const int s = surf % nsp;
    if (s < m.nx * m.ny) return 2;
    if (s < m.nx * m.ny + (m.nx + 1) * m.ny) return 0;
    return 1;
}

__device__ __forceinline__ void neigh_cells(const Grid &m, int surf, int &c1, int &c2)
{
    const int nx = m.nx, ny = m.ny;
    const int nsp = nx * ny + (nx + 1) * ny + (ny + 1) * nx;
    const int n = surf_normal(m, surf);
    const int iz = surf / nsp;
    int s = surf - iz * nsp;
    if (n == 0) {
        const int iy = (s - nx * ny) / (nx + 1), ix = (s - nx * ny) % (nx + 1);
        c1 = ix > 0 ? iy * nx + ix - 1 : -1;
        c2 = ix < nx ? iy * nx + ix : -1;
    } else if (n == 1) {
        s -= nx * ny + (nx + 1) * ny;
        const int iy = s % (ny + 1), ix = s / (ny + 1);
        c1 = iy > 0 ? (iy - 1) * nx + ix : -1;
        c2 = iy < ny ? iy * nx + ix : -1;
    } else {
        const int iy = s / nx, ix = s % nx;
        c1 = iz > 0 ? iy * nx + ix : -1;  // one plane: the top faces (iz = 1) have the cell below
        c2 = iz < 1 ? iy * nx + ix : -1;
    }
}

__device__ __forceinline__ int coarse_surf(const Grid &m, int cell, int s)
{
    const int nx = m.nx, ny = m.ny;
    const int ix = cell % nx, iy = cell / nx;
    const int nxy = (nx + 1) * ny + (ny + 1) * nx;
    switch (s) {
    case EAST: return (nx + 1) * iy + ix + 1 + nx * ny;
    case WEST: return (nx + 1) * iy + ix + nx * ny;
    case NORTH: return nx * ny + (nx + 1) * ny + (ny + 1) * ix + iy + 1;
    case SOUTH: return (nx + 1) * ny + (ny + 1) * ix + iy + nx * ny;
    case BOTTOM: return nx * iy + ix;
    default: return nx * ny + nxy + nx * iy + ix;  // TOP
    }
}

__device__ __forceinline__ double area_cell(const Grid &m, int cell, int s)
{
    const int ix = cell % m.nx, iy = cell / m.nx;
    if (s == EAST || s == WEST) return m.dy[iy] * m.hz;
    if (s == NORTH || s == SOUTH) return m.dx[ix] * m.hz;
    return m.dx[ix] * m.dy[iy];
}

__device__ __forceinline__ double thickness(const Grid &m, int cell, int normal)
{
    if (normal == 0) return m.dx[cell % m.nx];
    if (normal == 1) return m.dy[cell / m.nx];
    return m.hz;
}

// --- set-up kernels --------------------------------------------------------------

// MoCSweeper::get_pin_flux_1g (PIN_PLANE, moc_sweeper.cpp:293-326) for all
// groups: item t = (pin, group). With negative_fixup the coarse flux is
// clamped at 0, as CMFD::solve does before homogenizing (cmfd.cpp:175-184).
__device__ __forceinline__ void pin_flux_item(const CmfdView &v, double *cflux, int t)
{
    if (t >= v.npin * v.G) return;
    const int ip = t / v.G, g = t % v.G;
    double vs = 0.0, pf = 0.0;
    for (int r = v.pin_off[ip]; r < v.pin_off[ip + 1]; r++) {
        vs += v.vol[r];
        pf += v.phi[(size_t)r * v.G + g] * v.vol[r];
    }
    const double c = pf / vs;
    cflux[(size_t)g * v.nc + v.pin_cell[ip]] = v.neg_fixup ? fmax(c, 0.0) : c;
}

__global__ void k_pin_flux_all(CmfdView v, double *cflux)
{
    pin_flux_item(v, cflux, blockIdx.x * blockDim.x + threadIdx.x);
}

// XSMeshHomogenized::homogenize_region_flux: flux-area weighted Sigma_t,
// nu Sigma_f, Sigma_f; chi weighted by the fission source; Sigma_s[g <- g']
// weighted by phi_g'. One thread per (pin, group). What the coarse solve
// reads per cell (nu Sigma_f, chi, Sigma_s, like the coarse flux) is stored
// cell-fastest, [g][c] and [g][g'][c], so the lanes of a wave, which hold
// consecutive cells, read consecutive addresses; Sigma_t, Sigma_f and the
// removal stay [c][g].
struct HomogArrays {
    double *htr, *hnf, *hf, *hch, *hrm, *hsc;
};

__device__ __forceinline__ void homogenize_item(const CmfdView &v, const HomogArrays &h, int t)
{
    if (t >= v.npin * v.G) return;
    double *htr = h.htr, *hnf = h.hnf, *hf = h.hf, *hch = h.hch, *hrm = h.hrm, *hsc = h.hsc;
    const int G = v.G;
    const int ip = t / G, g = t % G;
    const int cell = v.pin_cell[ip];
    const int r0 = v.pin_off[ip], r1 = v.pin_off[ip + 1];
    double fs_sum = 0.0, ch = 0.0, fvs = 0.0, tr = 0.0, nf = 0.0, fi = 0.0;
    for (int r = r0; r < r1; r++) {
        const int m = v.fsr_mat[r];
        const double a = v.area[r];
        double fsv = 0.0;
        for (int gg = 0; gg < G; gg++) fsv += v.mat_nf[m * G + gg] * v.phi[(size_t)r * G + gg] * a;
        fs_sum += fsv;
        ch += fsv * v.mat_ch[m * G + g];
        const double fl = v.phi[(size_t)r * G + g];
        fvs += a * fl;
        tr += a * fl * v.mat_tr[m * G + g];
        nf += a * fl * v.mat_nf[m * G + g];
        fi += a * fl * v.mat_f[m * G + g];
    }
    const double htr_g = tr / fvs;
    htr[(size_t)cell * G + g] = htr_g;
    hnf[(size_t)g * v.nc + cell] = nf / fvs;
    hf[(size_t)cell * G + g] = fi / fvs;
    hch[(size_t)g * v.nc + cell] = (fs_sum > 0.0) ? ch / fs_sum : ch;
    double self = 0.0;
    for (int gg = 0; gg < G; gg++) {
        double sc = 0.0, ss = 0.0;
        for (int r = r0; r < r1; r++) {
            const int m = v.fsr_mat[r];
            const double a = v.area[r];
            const double fgg = v.phi[(size_t)r * G + gg];
            ss += fgg * a;
            sc += v.mat_sc[((size_t)m * G + g) * G + gg] * a * fgg;
        }
        if (sc > 0.0) sc /= ss;
        hsc[((size_t)g * G + gg) * v.nc + cell] = sc;
        if (gg == g) self = sc;
    }
    hrm[(size_t)cell * G + g] = htr_g - self;
}

__global__ void k_homogenize(CmfdView v, HomogArrays h)
{
    homogenize_item(v, h, blockIdx.x * blockDim.x + threadIdx.x);
}

struct CoupleArgs {
    Grid m;
    int G, n_surf, has_radial;
    const int32_t *radial_flag;  // if set, read instead of has_radial (batch graphs)
    int bc_lo[3], bc_hi[3];  // [normal]: low-side and high-side boundary kinds
    const double *htr, *cflux, *cur, *sfl;
    double *dtil, *stil, *dhat, *shat;
    double *part, *part_old;  // partial currents [n_surf][G][2] (null: not kept)
};

// CMFD::setup_solve, surface couplings (cmfd.cpp:423-490): item t = (surface, group)
__device__ __forceinline__ void couple_item(const CoupleArgs &a, int t)
{
    if (t >= a.n_surf * a.G) return;
    const int has_radial = a.radial_flag ? *a.radial_flag : a.has_radial;
    const int is = t / a.G, g = t % a.G, G = a.G;
    int c1, c2;
    neigh_cells(a.m, is, c1, c2);
    const int norm = surf_normal(a.m, is);
    auto bdiff = [](int kind) { return kind == MOCGPU_BC_VACUUM ? 0.5 / 2.0 : 0.0 / 2.0; };
    const double d1 = c1 > -1 ? (1.0 / (3.0 * a.htr[(size_t)c1 * G + g])) / thickness(a.m, c1, norm) : bdiff(a.bc_lo[norm]);
    const double d2 = c2 > -1 ? (1.0 / (3.0 * a.htr[(size_t)c2 * G + g])) / thickness(a.m, c2, norm) : bdiff(a.bc_hi[norm]);
    const double dt = 2.0 * d1 * d2 / (d1 + d2);
    const double st = (d1 > 0.0) ? d1 / (d1 + d2) : d2 / (d1 + d2);
    a.dtil[(size_t)is * G + g] = dt;
    a.stil[(size_t)is * G + g] = st;
    double dh = 0.0, sh = 0.0;
    if (norm != 2 && has_radial) {
        const double j = a.cur[(size_t)is * G + g];
        const double sf = a.sfl[(size_t)is * G + g];
        const size_t cg = (size_t)g * a.m.nx * a.m.ny;
        const double fl = c1 >= 0 ? a.cflux[cg + c1] : 0.0;
        const double fr = c2 >= 0 ? a.cflux[cg + c2] : 0.0;
        dh = (j + dt * (fr - fl)) / (fl + fr);
        if (!isfinite(dh)) dh = 0.0;
        sh = (c1 >= 0) ? (sf - st * fl - (1.0 - st) * fr) / (fl + fr) : (sf - st * fr) / fr;
    }
    a.dhat[(size_t)is * G + g] = dh;
    a.shat[(size_t)is * G + g] = sh;
}

__global__ void k_couple(CoupleArgs a)
{
    couple_item(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// CMFD::setup_solve, matrix rows (cmfd.cpp:506-544): item t = (group, cell)
__device__ __forceinline__ void assemble_item(const Grid &m, int G, const double *hrm, const double *dtil,
                                              const double *dhat, double *coef, int t)
{
    const int nc = m.nx * m.ny;
    if (t >= nc * G) return;
    const int g = t / nc, c = t % nc;
    const int ix = c % m.nx, iy = c / m.nx;
    double diag = m.dx[ix] * m.dy[iy] * m.hz * hrm[(size_t)c * G + g];
    double off[4] = {0.0, 0.0, 0.0, 0.0};  // east, north, west, south
    const int order[6] = {EAST, NORTH, WEST, SOUTH, TOP, BOTTOM};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int s = order[k];
        const int surf = coarse_surf(m, c, s);
        const double ar = area_cell(m, c, s);
        double dh = dhat[(size_t)surf * G + g];
        if (s == WEST || s == SOUTH || s == BOTTOM) dh = -dh;
        const double dt = dtil[(size_t)surf * G + g];
        diag += ar * (dt + dh);
        if (k < 4) off[k] = ar * (dh - dt);
    }
    const size_t pl = (size_t)G * nc;
    const size_t i = (size_t)g * nc + c;
    coef[0 * pl + i] = iy > 0 ? off[3] : 0.0;         // south
    coef[1 * pl + i] = ix > 0 ? off[2] : 0.0;         // west
    coef[2 * pl + i] = diag;
    coef[3 * pl + i] = ix < m.nx - 1 ? off[0] : 0.0;  // east
    coef[4 * pl + i] = iy < m.ny - 1 ? off[1] : 0.0;  // north
}

__global__ void k_assemble(Grid m, int G, const double *hrm, const double *dtil, const double *dhat, double *coef)
{
    assemble_item(m, G, hrm, dtil, dhat, coef, blockIdx.x * blockDim.x + threadIdx.x);
}

// CMFD::store_currents (cmfd.cpp:555-638): the net current and surface flux
// the coarse solution implies.
__device__ __forceinline__ void store_currents_item(const CoupleArgs &a, double *cur, double *sfl, int t)
{
    if (t >= a.n_surf * a.G) return;
    const int is = t / a.G, g = t % a.G, G = a.G;
    int c1, c2;
    neigh_cells(a.m, is, c1, c2);
    const size_t cg = (size_t)g * a.m.nx * a.m.ny;
    const double fr = c2 >= 0 ? a.cflux[cg + c2] : 0.0;
    const double fl = c1 >= 0 ? a.cflux[cg + c1] : 0.0;
    const double dh = a.dhat[(size_t)is * G + g], dt = a.dtil[(size_t)is * G + g];
    const double j = -dt * (fr - fl) + dh * (fr + fl);
    const double sh = a.shat[(size_t)is * G + g], st = a.stil[(size_t)is * G + g];
    const double sf = (c1 >= 0) ? st * fl + (1.0 - st) * fr + sh * (fl + fr) : st * fr + sh * (fl + fr);
    sfl[(size_t)is * G + g] = sf;
    cur[(size_t)is * G + g] = j;
    if (a.part) {
        // CMFD::store_currents (cmfd.cpp:560-599): the previous partial
        // currents become the old ones, then (sf/4 + J/2, sf/4 - J/2)
        const size_t k = ((size_t)is * G + g) * 2;
        a.part_old[k] = a.part[k];
        a.part_old[k + 1] = a.part[k + 1];
        a.part[k] = 0.25 * sf + 0.5 * j;
        a.part[k + 1] = 0.25 * sf - 0.5 * j;
    }
}

__global__ void k_store_currents(CoupleArgs a, double *cur, double *sfl)
{
    store_currents_item(a, cur, sfl, blockIdx.x * blockDim.x + threadIdx.x);
}

// MoCSweeper::set_pin_flux_1g (PIN_PLANE, moc_sweeper.cpp:360-433) for all
// groups; a non-positive coarse flux raises the flag (the reference throws).
__device__ __forceinline__ void set_pin_flux_item(const CmfdView &v, const double *cflux, int *bad, int t)
{
    if (t >= v.npin * v.G) return;
    const int ip = t / v.G, g = t % v.G;
    const double pn = cflux[(size_t)g * v.nc + v.pin_cell[ip]];
    if (!(pn > 0.0)) {
        atomicOr(bad, 1);
        return;
    }
    double fm = 0.0;
    for (int r = v.pin_off[ip]; r < v.pin_off[ip + 1]; r++) fm += v.phi[(size_t)r * v.G + g] * v.area[r];
    fm /= v.pin_area[ip];
    const double f = pn / fm;
    for (int r = v.pin_off[ip]; r < v.pin_off[ip + 1]; r++) v.phi[(size_t)r * v.G + g] *= f;
}

__global__ void k_set_pin_flux_all(CmfdView v, const double *cflux, int *bad)
{
    set_pin_flux_item(v, cflux, bad, blockIdx.x * blockDim.x + threadIdx.x);
}

// info[0]: coarse power iterations, info[1]: non-positive flux flag of the
// rescale, info[2]: BiCGSTAB iterations summed over the solve
__global__ void k_zero_info(int *info)
{
    if (threadIdx.x < 3) info[threadIdx.x] = 0;
}

// --- the coarse power iteration ------------------------------------------------

struct SolveArgs {
    int nc, nx, G, max_iter;
    double k_tol, psi_tol, rr;
    const double *coef;  // [5][G][nc]
    const double *hnf, *hch, *hsc, *cvol;
    const uint64_t *smask;  // [G]: the groups g' != g that scatter into g
    double *cflux;  // [G][nc]
    double *gscr;   // [3][nc] global scratch of the solves that keep the rows in LDS
    double *k;
    int *info;
};

// One DPP lane move of a double (two 32-bit moves); lanes the control leaves
// without a source, or whose row is masked off, read 0.
template <int CTRL, int ROWS = 0xf> __device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave by DPP moves (VALU latency, no LDS trip):
// pairs and quads (quad_perm), half rows and rows (mirrors), then rows 0+1
// and 2+3 (row_bcast:15) and the two halves (row_bcast:31), which leaves the
// total in lane 63. Every lane returns it.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xb1>(v);        // quad_perm [1,0,3,2]
    v += dpp_f64<0x4e>(v);        // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);       // row_half_mirror
    v += dpp_f64<0x140>(v);       // row_mirror
    v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Workgroup scratch of the coarse solve. Reductions and matrix-vector
// products each use one of two LDS banks in turn, so each needs a single
// barrier: a bank is rewritten only at the second use after, and every lane
// has finished reading it before it reaches the barrier of the use between.
// All lanes make the same sequence of calls, so the bank indices agree.
// The solve runs in workgroups of NT lanes (NT / 64 waves). NT is 512 for
// every mesh: with more lanes the per-lane register budget (4 cells' rows and
// BiCGSTAB vectors need about 240 VGPRs) is exceeded and the solve spills.
template <int NT> constexpr int red_bank() { return 2 * (NT / 64); }  // doubles per reduction bank

template <int NT> struct Scratch {
    double *red;  // [2][red_bank<NT>()] wave partial sums
    double *buf;  // [2][n] operand of the matrix-vector products
    int n;        // entries per bank (K * NT)
    int rb, bb;   // bank of the next reduction / product
};

// Sum over the workgroup; every lane returns the same value (lane 0 of each
// wave publishes, all lanes add the wave sums in wave order).
template <int NT> __device__ __forceinline__ double block_sum(double v, Scratch<NT> &S)
{
    constexpr int NW = NT / 64;
    v = wave_sum(v);
    double *red = S.red + S.rb * red_bank<NT>();
    S.rb ^= 1;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++) s += red[w];
    return s;
}

template <int NT> __device__ __forceinline__ void block_sum2(double a, double b, Scratch<NT> &S, double &sa, double &sb)
{
    constexpr int NW = NT / 64;
    a = wave_sum(a);
    b = wave_sum(b);
    double *red = S.red + S.rb * red_bank<NT>();
    S.rb ^= 1;
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = a;
        red[NW + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    sa = 0.0;
    sb = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        sa += red[w];
        sb += red[NW + w];
    }
}

// A lane's K cell slots: c = threadIdx.x + k * NT. Slots past the mesh (c >=
// nc) are dummy cells of the LDS vectors (K * NT entries): their rows are
// zero and reach no neighbour, and their vector entries stay exact zeros
// (inverse diagonal 1), so the BiCGSTAB loop runs without per-slot branches
// and the dummies add nothing to any sum.
template <int K> struct Lane {
    int c[K];   // cells of this lane: threadIdx.x + k * NT
    int ix[K];  // their x index
    bool on[K]; // c < nc (a mesh cell)
};

// Up to 3 cells per lane the loop computes the dummy slots too (no branches);
// with more, per-slot branches skip them: without those, the compiler keeps
// enough values live to spill at 4 cells per lane and beyond.
template <int K> constexpr bool branch_free() { return K <= 3; }

// The matrix of one group for this lane's cells. Up to 4 cells per lane the
// five coefficients of each row stay in registers for the whole group solve;
// with more, every product reloads them from cache (the registers would not
// hold them beside the BiCGSTAB vectors).
constexpr int kRegCoefMaxK = 4;

template <int K> struct Rows {
    double a[K <= kRegCoefMaxK ? 5 : 1][K <= kRegCoefMaxK ? K : 1];  // south, west, diagonal, east, north
    const double *cf;  // coefficients of group g (reloaded when not held)
    double *rl;        // [5][nc] the rows in LDS (solves with RL), else null
};

// Which of the five row entries (south, west, diagonal, east, north) of cell
// c exist, and the cell each one multiplies (c itself where the neighbour is
// missing, so that the product reads in bounds and adds an exact 0).
__device__ __forceinline__ bool row_has(const SolveArgs &A, int c, int ix, int j)
{
    switch (j) {
    case 0: return c >= A.nx;
    case 1: return ix > 0;
    case 3: return ix < A.nx - 1;
    case 4: return c + A.nx < A.nc;
    default: return true;
    }
}
__device__ __forceinline__ int row_col(const SolveArgs &A, int c, int ix, int j)
{
    const int off[5] = {-A.nx, -1, 0, 1, A.nx};
    return row_has(A, c, ix, j) ? c + off[j] : c;
}

// p plus an opaque zero: the compiler cannot hoist loads through it out of a
// loop (where the values would sit in registers that the loop's vectors need)
template <class T> __device__ __forceinline__ T *opaque(T *p)
{
    int zero = 0;
    __asm__ volatile("" : "+v"(zero));
    return p + zero;
}

template <int K, bool RL> __device__ __forceinline__ void load_rows(const SolveArgs &A, int g, const Lane<K> &L, Rows<K> &R)
{
    R.cf = A.coef + (size_t)g * A.nc;
    if constexpr (RL) {
        // each lane stages its own cells' rows; only it reads them
        const size_t pl = (size_t)A.G * A.nc;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (L.on[k])
#pragma unroll
                for (int j = 0; j < 5; j++)
                    R.rl[j * A.nc + L.c[k]] = row_has(A, L.c[k], L.ix[k], j) ? R.cf[j * pl + L.c[k]] : 0.0;
    } else if constexpr (K <= kRegCoefMaxK) {
        const size_t pl = (size_t)A.G * A.nc;
#pragma unroll
        for (int j = 0; j < 5; j++)
#pragma unroll
            for (int k = 0; k < K; k++)
                R.a[j][k] = (L.on[k] && row_has(A, L.c[k], L.ix[k], j)) ? R.cf[j * pl + L.c[k]] : 0.0;
    }
}

// v = A y for group g: y staged in an LDS bank, rows summed in sorted column
// order (south, west, diagonal, east, north) as a CSR product would. A
// missing neighbour enters as 0 times the cell's own entry, which leaves the
// sum unchanged, so the five reads of a row carry no branches and issue
// together. y and v may be the same array. Returns the bank, where y stays
// readable until the second product after this one (with RL, whose single
// bank each product overwrites after a barrier, until the next one).
template <int K, int NT, bool RL>
__device__ __forceinline__ const double *spmv(const SolveArgs &A, const Rows<K> &R, const Lane<K> &L,
                                              const double (&y)[K], double (&v)[K], Scratch<NT> &S)
{
    double *buf;
    if constexpr (RL) {
        // one operand bank: every lane has read the previous operand first
        buf = S.buf;
        __syncthreads();
    } else {
        buf = S.buf + S.bb * (K * NT);
        S.bb ^= 1;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
        if (branch_free<K>() || L.on[k]) buf[L.c[k]] = y[k];
    __syncthreads();
    const size_t pl = (size_t)A.G * A.nc;
    // every product reloads the rows of K > kRegCoefMaxK cells
    const double *cf = (K > kRegCoefMaxK && !RL) ? opaque(R.cf) : R.cf;
    const double *rl = (K > kRegCoefMaxK && RL) ? opaque(R.rl) : R.rl;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        if (!branch_free<K>() && !L.on[k]) {
            v[k] = 0.0;
            continue;
        }
        const int c = L.c[k];
        const int ix = L.ix[k];
        double a[5];
#pragma unroll
        for (int j = 0; j < 5; j++) {
            if constexpr (RL)
                a[j] = rl[j * A.nc + c];
            else if constexpr (K <= kRegCoefMaxK)
                a[j] = R.a[j][k];
            else
                a[j] = (L.on[k] && row_has(A, c, ix, j)) ? cf[j * pl + c] : 0.0;
        }
        double y5[5];
#pragma unroll
        for (int j = 0; j < 5; j++) y5[j] = buf[L.on[k] ? row_col(A, c, ix, j) : c];
#pragma unroll
        for (int j = 0; j < 5; j++) s = fma(a[j], y5[j], s);
        v[k] = s;
    }
    return buf;
}

// Device code is compiled without floating-point contraction, so that a
// problem's arithmetic is the same at any cells per lane (a batch runs every
// problem at those of its largest mesh): the hot loops below write their
// fused multiply-adds explicitly.
//
// Jacobi-preconditioned BiCGSTAB with the relative test |r| <= tol |b| and a
// restart when rho vanishes (van der Vorst 1992, the iteration of Eigen's
// BiCGSTAB<SparseMatrix<double>> that cmfd.cpp:291 calls). The preconditioned
// directions y = M^-1 p and z = M^-1 s are read back from the product's LDS
// operand, so x is updated as (x + alpha y) + w z. The loop-invariant vectors
// (the inverse diagonal, the shadow residual r0 and b) live in LDS slots that
// only the lane owning the cell touches, which keeps the rest in registers.
template <int K, int NT, bool RL = false>
__device__ int bicgstab(const SolveArgs &A, const Rows<K> &R, const Lane<K> &L, double (&x)[K], double tol,
                        int max_iter, Scratch<NT> &S, double *sinv, double *sr0, const double *sb)
{
    const size_t pl = (size_t)A.G * A.nc;
    double r[K], p[K], v[K], t[K];
    // the Jacobi preconditioner: the inverse diagonal (in LDS, or with the
    // rows in LDS in the global scratch, reloaded every iteration)
    const double *si = sinv;
    auto inv_diag = [&](int k) -> double { return si[L.c[k]]; };
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (!branch_free<K>() && !L.on[k]) continue;
        const double d = L.on[k] ? R.cf[2 * pl + L.c[k]] : 0.0;
        sinv[L.c[k]] = (d != 0.0) ? 1.0 / d : 1.0;
    }
    spmv<K, NT, RL>(A, R, L, x, t, S);
    double lr0 = 0.0, lb = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double bk = (branch_free<K>() || L.on[k]) ? sb[L.c[k]] : 0.0;
        r[k] = bk - t[k];
        if (branch_free<K>() || L.on[k]) sr0[L.c[k]] = r[k];
        p[k] = 0.0;
        v[k] = 0.0;
        lr0 = fma(r[k], r[k], lr0);
        lb = fma(bk, bk, lb);
    }
    double r0_sq, rhs_sq;
    block_sum2(lr0, lb, S, r0_sq, rhs_sq);
    if (rhs_sq == 0.0) {
#pragma unroll
        for (int k = 0; k < K; k++) x[k] = 0.0;
        return 0;
    }
    double rho = 1.0, alpha = 1.0, w = 1.0;
    const double tol2 = tol * tol * rhs_sq;
    const double eps = 2.220446049250313e-16;
    const double eps2 = eps * eps;
    int i = 0, restarts = 0;
    while (true) {
        if constexpr (RL) {
            si = opaque(sinv);
            sb = opaque(sb);
        }
        double lrr = 0.0, lrho = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            lrr = fma(r[k], r[k], lrr);
            if (branch_free<K>() || L.on[k]) lrho = fma(sr0[L.c[k]], r[k], lrho);
        }
        double rr, rho_new;
        block_sum2(lrr, lrho, S, rr, rho_new);
        if (!(rr > tol2 && i < max_iter)) break;
        const double rho_old = rho;
        rho = rho_new;
        if (fabs(rho) < eps2 * r0_sq) {
            spmv<K, NT, RL>(A, R, L, x, t, S);
            double l = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                r[k] = ((branch_free<K>() || L.on[k]) ? sb[L.c[k]] : 0.0) - t[k];
                if (branch_free<K>() || L.on[k]) sr0[L.c[k]] = r[k];
                l = fma(r[k], r[k], l);
            }
            rho = r0_sq = block_sum(l, S);
            if (restarts++ == 0) i = 0;
        }
        const double beta = (rho / rho_old) * (alpha / w);
#pragma unroll
        for (int k = 0; k < K; k++) {
            p[k] = fma(beta, fma(-w, v[k], p[k]), r[k]);
            t[k] = ((branch_free<K>() || L.on[k]) ? inv_diag(k) : 1.0) * p[k];  // y
        }
        const double *yb = spmv<K, NT, RL>(A, R, L, t, v, S);
        double l = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (branch_free<K>() || L.on[k]) l = fma(sr0[L.c[k]], v[k], l);
        alpha = rho / block_sum(l, S);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (branch_free<K>() || L.on[k]) x[k] = fma(alpha, yb[L.c[k]], x[k]);
            r[k] = fma(-alpha, v[k], r[k]);                 // s
            t[k] = ((branch_free<K>() || L.on[k]) ? inv_diag(k) : 1.0) * r[k];  // z
        }
        const double *zb = spmv<K, NT, RL>(A, R, L, t, t, S);  // t = A z
        double ltt = 0.0, lts = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            ltt = fma(t[k], t[k], ltt);
            lts = fma(t[k], r[k], lts);
        }
        double tt, ts;
        block_sum2(ltt, lts, S, tt, ts);
        w = (tt > 0.0) ? ts / tt : 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (branch_free<K>() || L.on[k]) x[k] = fma(w, zb[L.c[k]], x[k]);
            r[k] = fma(-w, t[k], r[k]);
        }
        ++i;
    }
    return i;
}

// The per-cell sums below read all K cells of one step together (a dummy
// slot reads cell 0 and its value is dropped), so a lane has K loads in
// flight instead of one; each cell's sum keeps its order of terms. The
// arrays are cell-fastest ([g][c], [g][g'][c]): a wave's lanes, which hold
// consecutive cells, read consecutive addresses.
// 1 at a lane's real cells, 0 at its dummies. Values loaded at clamped cells
// are multiplied by it instead of selected: a select lets the compiler move
// the load under a branch of its own, where it waits alone, while loads used
// unconditionally are issued together. Every factor is finite, so a dummy
// gets a zero (of either sign: it adds nothing to any sum) and a real cell
// its value unchanged.
template <int K> __device__ __forceinline__ double on_factor(const Lane<K> &L, int k) { return L.on[k] ? 1.0 : 0.0; }

template <int K> __device__ __forceinline__ void lane_cells(const Lane<K> &L, int (&ci)[K])
{
#pragma unroll
    for (int k = 0; k < K; k++) ci[k] = L.on[k] ? L.c[k] : 0;
}

// CMFD source of group g: V (chi fs + sum_{g' != g} Sigma_s[g <- g'] phi_g')
// (cmfd.cpp:225-231), from this lane's own cells.
template <int K>
__device__ __forceinline__ void cmfd_source(const SolveArgs &A, int g, const Lane<K> &L, const double (&fsc)[K],
                                            double (&b)[K])
{
    const int G = A.G;
    const size_t nc = (size_t)A.nc;
    int ci[K];
    lane_cells<K>(L, ci);
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = A.hch[g * nc + ci[k]] * fsc[k];
    // the groups g' no material scatters from into g add exact zeros: they
    // are skipped, the others added in increasing g'
    for (uint64_t m = A.smask[g]; m != 0; m &= m - 1) {
        const int gg = __builtin_ctzll(m);
        const double *sc = A.hsc + ((size_t)g * G + gg) * nc, *fl = A.cflux + gg * nc;
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += sc[ci[k]] * fl[ci[k]];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        b[k] = s[k] * A.cvol[ci[k]] * on_factor<K>(L, k);
    }
}

template <int K>
__device__ __forceinline__ double fission_source(const SolveArgs &A, const Lane<K> &L, double k_eff, double (&fsc)[K])
{
    const double rk = 1.0 / k_eff;
    const size_t nc = (size_t)A.nc;
    int ci[K];
    lane_cells<K>(L, ci);
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int g = 0; g < A.G; g++) {
        const double *nf = A.hnf + g * nc, *fl = A.cflux + g * nc;
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += nf[ci[k]] * fl[ci[k]];
    }
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double sk = L.on[k] ? s[k] : 0.0;
        tot += sk;
        fsc[k] = sk * rk;
    }
    return tot;
}

// This lane's share of the coarse fission total sum_c sum_g nu-Sigma_f phi,
// summed term by term in cell order, then group order.
template <int K> __device__ __forceinline__ double lane_fission_total(const SolveArgs &A, const Lane<K> &L)
{
    const size_t nc = (size_t)A.nc;
    int ci[K];
    lane_cells<K>(L, ci);
    double lf = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
#pragma unroll 8
        for (int g = 0; g < A.G; g++) {
            const double v = A.hnf[g * nc + ci[k]] * A.cflux[g * nc + ci[k]];
            lf += L.on[k] ? v : 0.0;
        }
    }
    return lf;
}

// This lane's share of the squared change of the coarse fission source.
template <int K>
__device__ __forceinline__ double lane_source_change(const Lane<K> &L, const double (&fsc)[K], const double *fs_old)
{
    double le = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double d = fsc[k] - fs_old[L.on[k] ? L.c[k] : 0];
        le += L.on[k] ? d * d : 0.0;
    }
    return le;
}

// ||A x - b||^2 of group g with x the stored coarse flux (cmfd.cpp:640-666)
template <int K, int NT, bool RL = false>
__device__ __forceinline__ double residual2(const SolveArgs &A, int g, const Rows<K> &R, const Lane<K> &L,
                                            const double (&b)[K], double (&x)[K], double (&t)[K], Scratch<NT> &S)
{
#pragma unroll
    for (int k = 0; k < K; k++) x[k] = L.on[k] ? A.cflux[(size_t)g * A.nc + L.c[k]] : 0.0;
    spmv<K, NT, RL>(A, R, L, x, t, S);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double d = t[k] - b[k];
        s = fma(d, d, s);
    }
    return s;
}

// CMFD::solve (cmfd.cpp:171-279) by one workgroup of NT lanes, K cells per
// lane; red holds 2 * red_bank<NT>() doubles. lds holds kCmfdLdsVectors *
// K * NT doubles, or with RL (rows in LDS) 7 * nc: one
// product operand, the shadow residual and the group's five rows, while the
// previous fission source, the right-hand side and the inverse diagonal go to
// A.gscr. The launches use K = 1 or 2 without RL (meshes of up to 1024
// cells; larger ones run the wide solve); the K > 2 and RL forms are the
// general form of this solve and are not launched.
template <int K, int NT, bool RL = false>
__device__ __forceinline__ void cmfd_solve(const SolveArgs &A, double *lds, double *red)
{
    // the nc-long vectors hold no dummy slots: every access must skip them
    static_assert(!RL || !branch_free<K>(), "rows in LDS need per-slot branches");
    constexpr int KN = K * NT;  // entries of each LDS vector (cells and dummies)
    Scratch<NT> S{red, lds, KN, 0, 0};
    double *fs_old, *sinv, *sr0, *sb;
    Rows<K> R;
    if constexpr (RL) {
        sr0 = lds + A.nc;         // [nc] BiCGSTAB: shadow residual
        R.rl = lds + 2 * A.nc;    // [5][nc] the group's rows
        fs_old = A.gscr;          // [nc] previous coarse fission source
        sb = A.gscr + A.nc;       // [nc] BiCGSTAB: right-hand side
        sinv = A.gscr + 2 * A.nc; // [nc] BiCGSTAB: inverse diagonal
    } else {
        fs_old = lds + 2 * KN;    // [KN] previous coarse fission source
        sinv = lds + 3 * KN;      // [KN] BiCGSTAB: inverse diagonal
        sr0 = lds + 4 * KN;       // [KN] BiCGSTAB: shadow residual
        sb = lds + 5 * KN;        // [KN] BiCGSTAB: right-hand side
        R.rl = nullptr;
    }
    Lane<K> L;
#pragma unroll
    for (int k = 0; k < K; k++) {
        L.c[k] = threadIdx.x + k * NT;
        L.ix[k] = L.c[k] % A.nx;
        L.on[k] = L.c[k] < A.nc;
    }
    const int G = A.G;
    double k_eff = A.k[0];
    double fsc[K], b[K], x[K], t[K];
    double tfis = block_sum(fission_source<K>(A, L, k_eff, fsc), S);
    double r0 = 0.0;
    for (int g = 0; g < G; g++) {
        cmfd_source<K>(A, g, L, fsc, b);
        load_rows<K, RL>(A, g, L, R);
        r0 += block_sum(residual2<K, NT, RL>(A, g, R, L, b, x, t, S), S);
    }
    r0 = sqrt(r0) / (G * A.nc);
    const double tol = A.rr * r0;
    int iter = 0, n_bicg = 0;
    while (true) {
        iter++;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (L.on[k]) fs_old[L.c[k]] = fsc[k];
        fission_source<K>(A, L, k_eff, fsc);
        const double tfis_old = tfis;
        double ri = 0.0;
        for (int g = 0; g < G; g++) {
            cmfd_source<K>(A, g, L, fsc, b);
            load_rows<K, RL>(A, g, L, R);
            ri += block_sum(residual2<K, NT, RL>(A, g, R, L, b, x, t, S), S);
#pragma unroll
            for (int k = 0; k < K; k++)
                if (branch_free<K>() || L.on[k]) sb[L.c[k]] = b[k];  // 0 at dummies
            n_bicg += bicgstab<K, NT, RL>(A, R, L, x, tol, 150, S, sinv, sr0, sb);
#pragma unroll
            for (int k = 0; k < K; k++)
                if (L.on[k]) A.cflux[(size_t)g * A.nc + L.c[k]] = x[k];
        }
        ri = sqrt(ri) / (A.nc * G);
        tfis = block_sum(lane_fission_total<K>(A, L), S);
        const double k_old = k_eff;
        k_eff = k_eff * tfis / tfis_old;
        const double psi_err = sqrt(block_sum(lane_source_change<K>(L, fsc, fs_old), S));
        if (((fabs(k_eff - k_old) < A.k_tol) && (psi_err < A.psi_tol) && (ri / r0 < A.rr)) || (iter > A.max_iter))
            break;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
        if (L.on[k])
            for (int g = 0; g < G; g++) {
                double &f = A.cflux[(size_t)g * A.nc + L.c[k]];
                if (f < 0.0) f = -f;
            }
    if (threadIdx.x == 0) {
        A.k[0] = k_eff;
        A.info[0] = iter;
        A.info[2] = n_bicg;
    }
}

template <int K, int NT, bool RL = false> __global__ __launch_bounds__(NT) void k_cmfd_solve(SolveArgs A)
{
    extern __shared__ double lds[];
    __shared__ double red[2 * red_bank<NT>()];
    cmfd_solve<K, NT, RL>(A, lds, red);
}

// --- the coarse solve of meshes over 1024 cells -------------------------------
//
// One workgroup of NT lanes (NT / 64 waves), K cells per lane: lane l owns
// cells l + k * NT. What a lane keeps where:
//   * registers: the BiCGSTAB vectors x, r, p, v, t, the inverse diagonal and
//     the shadow residual of its cells, their coarse fission source and, on
//     512 lanes, their five row entries;
//   * LDS: [nx pad | operand bank 0 | operand bank 1 | nx pad];
//   * global scratch, touched once per group solve or power iteration: the
//     right-hand side and the previous coarse fission source.
// What bounds an iteration here is LDS traffic and the chain of barriers
// (three workgroup sums and two products), so a product reads only the
// north and south neighbours from LDS; the west and east neighbours (lanes
// l - 1 and l + 1) move across the wave by DPP, and only a wave's first and
// last lanes read theirs from LDS. Slots past the mesh are dummy cells with
// zero rows, zero vector entries and inverse diagonal 1, so no step branches on
// them. The pads hold zeros, so every neighbour read stays inside the
// allocation without clamps: a missing neighbour has a zero coefficient
// (assemble_item writes it so), and what it multiplies -- a pad zero or a
// finite entry of the other bank or of another cell -- adds an exact zero.
// The five products of a row are summed in the order south, west, diagonal,
// east, north, as in the 512-lane solve above, so every row's product is the
// same there and here; only the partition of the dot products differs.
// Where a lane keeps its cells' rows: in registers with 512 lanes (up to 6
// cells per lane: the 256-VGPR budget of two waves per SIMD holds them beside
// the vectors), or nowhere with 1024 lanes (every product reads them from
// cache).
template <int NT, int K> constexpr bool wide_rows_in_regs() { return NT <= 512 && K <= 6; }

template <int NT, int K> struct WideRows {
    double a[wide_rows_in_regs<NT, K>() ? 5 : 1][wide_rows_in_regs<NT, K>() ? K : 1];  // south, west, diagonal, east, north
    const double *cf;  // coefficients of group g: entry j of cell c at cf[j * G * nc + c]
};

// A DPP move whose pattern gives every lane a source lane (quad_perm,
// row_half_mirror, row_mirror): no old value to initialize.
template <int CTRL> __device__ __forceinline__ double dpp_all(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Wave shifts by one lane (lane i gets lane i - 1, or i + 1); the lane left
// without a source (0, or 63) gets an undefined value that the caller replaces.
__device__ __forceinline__ double dpp_wave_shr1(double v) { return dpp_all<0x138>(v); }
__device__ __forceinline__ double dpp_wave_shl1(double v) { return dpp_all<0x130>(v); }

// The sum of a row of 16 lanes in every lane of the row.
__device__ __forceinline__ double row_sum16(double v)
{
    v += dpp_all<0xb1>(v);
    v += dpp_all<0x4e>(v);
    v += dpp_all<0x141>(v);
    v += dpp_all<0x140>(v);
    return v;
}
__device__ __forceinline__ double read_lane_f64(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Workgroup sums of up to five values at once. Each row of 16 lanes sums its
// lanes by DPP (every lane of the row gets the same sum); lane 0 of each row
// publishes the row sums, value j of row r at j P + r with P = NT / 16 rows,
// in one of two banks, alternating, so each sum needs one barrier. Every
// wave then reads value j into its row j (value 4 into a second register),
// where each lane adds P / 16 row sums in a fixed order and the row adds its
// 16 lanes by the same DPP tree: every lane of every wave gets the same
// values, independent of timing.
template <int NT> constexpr int wide_red_rows() { return NT / 16; }
template <int NT> constexpr int wide_red_bank() { return 5 * wide_red_rows<NT>(); }  // doubles per bank
template <int NT> struct WideScratch {
    double *red;  // [2][wide_red_bank<NT>()]
    double *buf;  // operand bank 0 (bank 1 follows it)
    int rb, bb;
};
template <int NT, int N>
__device__ __forceinline__ void wide_sum(const double (&v)[N], WideScratch<NT> &S, double (&out)[N])
{
    constexpr int P = wide_red_rows<NT>();
    static_assert(N >= 1 && N <= 5, "at most five values per sum");
    static_assert(P % 16 == 0 && P <= 64, "whole rows of row sums");
    double *red = S.red + S.rb * wide_red_bank<NT>();
    S.rb ^= 1;
    double rs[N];
#pragma unroll
    for (int j = 0; j < N; j++) rs[j] = row_sum16(v[j]);
    if ((threadIdx.x & 15) == 0) {
        const int r = threadIdx.x >> 4;
#pragma unroll
        for (int j = 0; j < N; j++) red[j * P + r] = rs[j];
    }
    __syncthreads();
    const int l = threadIdx.x & 63, i = l & 15;
    constexpr int N4 = N < 4 ? N : 4;
    {
        const int j = (l >> 4) < N4 ? (l >> 4) : N4 - 1;  // rows past the last value repeat it
        const double *q = red + j * P + i;
        double t = q[0];
#pragma unroll
        for (int m = 1; m < P / 16; m++) t += q[16 * m];
        t = row_sum16(t);
#pragma unroll
        for (int j2 = 0; j2 < N4; j2++) out[j2] = read_lane_f64(t, 16 * j2);
    }
    if constexpr (N == 5) {
        const double *q = red + 4 * P + i;
        double t = q[0];
#pragma unroll
        for (int m = 1; m < P / 16; m++) t += q[16 * m];
        t = row_sum16(t);
        out[4] = read_lane_f64(t, 0);
    }
}
template <int NT> __device__ __forceinline__ double wide_sum1(double v, WideScratch<NT> &S)
{
    const double in[1] = {v};
    double out[1];
    wide_sum<NT, 1>(in, S, out);
    return out[0];
}

// The lane index through an opaque move: addresses formed from it are formed
// where they are used, not hoisted out of the solve loop (where they would hold
// registers that the BiCGSTAB vectors need).
__device__ __forceinline__ int opaque_lane()
{
    int l = threadIdx.x;
    __asm__ volatile("" : "+v"(l));
    return l;
}

// The group's rows for this lane's cells, and their inverse diagonal (1 where
// the diagonal is 0, as at the dummies).
template <int NT, int K>
__device__ __forceinline__ void wide_load_rows(const SolveArgs &A, int g, const Lane<K> &L, WideRows<NT, K> &R,
                                               double (&inv)[K])
{
    const size_t pl = (size_t)A.G * A.nc;
    R.cf = A.coef + (size_t)g * A.nc;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int cl = L.on[k] ? L.c[k] : 0;
        const double m = on_factor<K>(L, k);
        const double dg = R.cf[2 * pl + cl] * m;
        inv[k] = (dg != 0.0) ? 1.0 / dg : 1.0;
        if constexpr (wide_rows_in_regs<NT, K>()) {
            R.a[0][k] = R.cf[0 * pl + cl] * m;
            R.a[1][k] = R.cf[1 * pl + cl] * m;
            R.a[2][k] = dg;
            R.a[3][k] = R.cf[3 * pl + cl] * m;
            R.a[4][k] = R.cf[4 * pl + cl] * m;
        }
    }
}

// The stored coarse flux of group g at this lane's cells (0 at the dummies).
template <int K>
__device__ __forceinline__ void wide_load_flux(const SolveArgs &A, int g, const Lane<K> &L, double (&x)[K])
{
    const double *fl = A.cflux + (size_t)g * A.nc;
#pragma unroll
    for (int k = 0; k < K; k++) {
        x[k] = fl[L.on[k] ? L.c[k] : 0] * on_factor<K>(L, k);
    }
}

// v = A y for group g (y and v may be the same array); returns the operand
// bank, where y stays readable until the second product after this one.
template <int NT, int K>
__device__ __forceinline__ const double *wide_spmv(const SolveArgs &A, const WideRows<NT, K> &R, const Lane<K> &L,
                                                   const double (&y)[K], double (&v)[K], WideScratch<NT> &S)
{
    constexpr int KN = K * NT;
    double *buf = S.buf + S.bb * KN;
    S.bb ^= 1;
    const int l = opaque_lane();
    double *own = buf + l;
#pragma unroll
    for (int k = 0; k < K; k++) own[k * NT] = y[k];
    __syncthreads();
    const double *south = own - A.nx, *north = own + A.nx;
    const size_t pl = (size_t)A.G * A.nc;
    const int wl = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; k++) {
        // west and east neighbours (cells c - 1 and c + 1 are lanes l - 1 and
        // l + 1): moved across the wave by DPP, read from LDS only by the
        // wave's first and last lanes, whose neighbours sit in other waves
        double yw = dpp_wave_shr1(y[k]), ye = dpp_wave_shl1(y[k]);
        if (wl == 0) {
            __asm__ volatile("" ::: "memory");  // keeps the read in the branch
            yw = own[k * NT - 1];
        }
        if (wl == 63) {
            __asm__ volatile("" ::: "memory");
            ye = own[k * NT + 1];
        }
        double a0, a1, a2, a3, a4;
        if constexpr (wide_rows_in_regs<NT, K>()) {
            a0 = R.a[0][k];
            a1 = R.a[1][k];
            a2 = R.a[2][k];
            a3 = R.a[3][k];
            a4 = R.a[4][k];
        } else {
            const int c = l + k * NT;
            const bool on = c < A.nc;
            const double *cf = R.cf + (on ? c : 0);
            a0 = on ? cf[0 * pl] : 0.0;
            a1 = on ? cf[1 * pl] : 0.0;
            a2 = on ? cf[2 * pl] : 0.0;
            a3 = on ? cf[3 * pl] : 0.0;
            a4 = on ? cf[4 * pl] : 0.0;
        }
        double s = 0.0;
        s = fma(a0, south[k * NT], s);
        s = fma(a1, yw, s);
        s = fma(a2, y[k], s);
        s = fma(a3, ye, s);
        s = fma(a4, north[k * NT], s);
        v[k] = s;
    }
    return buf;
}

// The Jacobi-preconditioned BiCGSTAB of bicgstab() above, with the inverse
// diagonal and the shadow residual in registers and the right-hand side b in
// registers at the start (and in global scratch sb, read at a restart). ax is
// A x of the initial guess, which the caller has already formed; r0_sq
// returns ||b - A x||^2 of that guess.
//
// An iteration has four workgroup phases instead of five: the products
// v = A y and t = A z, alpha's sum, and one sum of five values after t, from
// which the next residual's norm and rho follow without a sum of their own,
// r = s - w t giving
//   r.r  = s.s - w t.s   (w = t.s / t.t),   r0.r = r0.s - w r0.t.
// Where the first would lose more than four digits to cancellation
// (r.r < 1e-4 s.s), or the second would (|r0.r| < 1e-4 |r0.s|), both are
// summed directly instead. Otherwise r.r carries a relative error of at most
// about 1e4 ulps, so the stopping test r.r > tol^2 b.b can differ from one on
// a direct sum (itself rounded in a different order) only where r.r lies
// within that error of the threshold.
template <int NT, int K>
__device__ int wide_bicgstab(const SolveArgs &A, const WideRows<NT, K> &R, const Lane<K> &L, const double (&inv)[K],
                             double (&x)[K], const double (&ax)[K], const double (&b)[K], double tol, int max_iter,
                             WideScratch<NT> &S, const double *sb, double &r0_sq)
{
    double r[K], p[K], v[K], t[K], r0[K];
    double ls[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gs[5];
#pragma unroll
    for (int k = 0; k < K; k++) {
        r[k] = b[k] - ax[k];
        r0[k] = r[k];
        p[k] = 0.0;
        v[k] = 0.0;
        ls[0] = fma(r[k], r[k], ls[0]);
        ls[1] = fma(b[k], b[k], ls[1]);
    }
    {
        const double in[2] = {ls[0], ls[1]};
        double out[2];
        wide_sum<NT, 2>(in, S, out);
        r0_sq = out[0];
        gs[1] = out[1];
    }
    const double rhs_sq = gs[1];
    if (rhs_sq == 0.0) {
#pragma unroll
        for (int k = 0; k < K; k++) x[k] = 0.0;
        return 0;
    }
    double rho = 1.0, alpha = 1.0, w = 1.0;
    const double tol2 = tol * tol * rhs_sq;
    const double eps = 2.220446049250313e-16;
    const double eps2 = eps * eps;
    double r0r0 = r0_sq;              // r0.r0 (changes at a restart)
    double rr = r0_sq, rho_new = r0_sq;  // r = r0 at the start
    int i = 0, restarts = 0;
    while (true) {
        if (!(rr > tol2 && i < max_iter)) break;
        const double rho_old = rho;
        rho = rho_new;
        if (fabs(rho) < eps2 * r0r0) {
            wide_spmv<NT, K>(A, R, L, x, t, S);
            const double *bl = sb + opaque_lane();
            double l = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                r[k] = bl[k * NT] - t[k];
                r0[k] = r[k];
                l = fma(r[k], r[k], l);
            }
            rho = r0r0 = wide_sum1<NT>(l, S);
            if (restarts++ == 0) i = 0;
        }
        // (rho / rho_old) (alpha / w) with one divide: one divide fewer on
        // the path from the last sum to the next product
        const double beta = (rho * alpha) / (rho_old * w);
#pragma unroll
        for (int k = 0; k < K; k++) {
            p[k] = fma(beta, fma(-w, v[k], p[k]), r[k]);
            t[k] = inv[k] * p[k];  // y
        }
        wide_spmv<NT, K>(A, R, L, t, v, S);  // v = A y; t keeps y
        double l = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) l = fma(r0[k], v[k], l);
        alpha = rho / wide_sum1<NT>(l, S);
#pragma unroll
        for (int k = 0; k < K; k++) {
            x[k] = fma(alpha, t[k], x[k]);
            r[k] = fma(-alpha, v[k], r[k]);  // s
            t[k] = inv[k] * r[k];            // z
        }
        const double *zb = wide_spmv<NT, K>(A, R, L, t, t, S) + opaque_lane();  // t = A z
#pragma unroll
        for (int j = 0; j < 5; j++) ls[j] = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            ls[0] = fma(t[k], t[k], ls[0]);
            ls[1] = fma(t[k], r[k], ls[1]);
            ls[2] = fma(r[k], r[k], ls[2]);
            ls[3] = fma(r0[k], r[k], ls[3]);
            ls[4] = fma(r0[k], t[k], ls[4]);
        }
        wide_sum<NT, 5>(ls, S, gs);
        const double tt = gs[0], ts = gs[1], ss = gs[2], r0s = gs[3], r0t = gs[4];
        w = (tt > 0.0) ? ts / tt : 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            x[k] = fma(w, zb[k * NT], x[k]);
            r[k] = fma(-w, t[k], r[k]);
        }
        rr = fma(-w, ts, ss);
        rho_new = fma(-w, r0t, r0s);
        if (!(rr > 1e-4 * ss && fabs(rho_new) > 1e-4 * fabs(r0s))) {
            double in[2] = {0.0, 0.0}, out[2];
#pragma unroll
            for (int k = 0; k < K; k++) {
                in[0] = fma(r[k], r[k], in[0]);
                in[1] = fma(r0[k], r[k], in[1]);
            }
            wide_sum<NT, 2>(in, S, out);
            rr = out[0];
            rho_new = out[1];
        }
        ++i;
    }
    return i;
}

// CMFD::solve (cmfd.cpp:171-279) as cmfd_solve() above, by NT lanes. lds
// holds nx + 2 K NT + nx doubles (the padded operand banks); A.gscr holds the
// previous coarse fission source and the right-hand side, K NT entries each.
template <int NT, int K> __device__ __forceinline__ void cmfd_solve_wide(const SolveArgs &A, double *lds, double *red)
{
    constexpr int KN = K * NT;
    WideScratch<NT> S{red, lds + A.nx, 0, 0};
    // The pads and both banks start as zeros: a product reads across the end
    // of its bank into the other one (the north neighbours of the last row,
    // the east neighbour of the last lane) before that bank is first written,
    // and LDS holds whatever the previous kernel left there (a NaN pattern
    // times a zero coefficient is NaN).
    for (int i = threadIdx.x; i < 2 * A.nx + 2 * KN; i += NT) lds[i] = 0.0;
    WideRows<NT, K> R;
    double *fs_old = A.gscr;
    double *sb = A.gscr + KN;
    Lane<K> L;
#pragma unroll
    for (int k = 0; k < K; k++) {
        L.c[k] = threadIdx.x + k * NT;
        L.ix[k] = 0;  // unused: the rows carry zeros for missing neighbours
        L.on[k] = L.c[k] < A.nc;
    }
    const int G = A.G;
    double k_eff = A.k[0];
    double fsc[K], b[K], x[K], t[K], inv[K];
    double tfis = wide_sum1<NT>(fission_source<K>(A, L, k_eff, fsc), S);
    double r0 = 0.0;
    for (int g = 0; g < G; g++) {
        cmfd_source<K>(A, g, L, fsc, b);
        wide_load_rows<NT, K>(A, g, L, R, inv);
        wide_load_flux<K>(A, g, L, x);
        wide_spmv<NT, K>(A, R, L, x, t, S);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double d = t[k] - b[k];
            s = fma(d, d, s);
        }
        r0 += wide_sum1<NT>(s, S);
    }
    r0 = sqrt(r0) / (G * A.nc);
    const double tol = A.rr * r0;
    int iter = 0, n_bicg = 0;
    while (true) {
        iter++;
#pragma unroll
        for (int k = 0; k < K; k++) fs_old[L.c[k]] = fsc[k];
        fission_source<K>(A, L, k_eff, fsc);
        const double tfis_old = tfis;
        double ri = 0.0;
        for (int g = 0; g < G; g++) {
            cmfd_source<K>(A, g, L, fsc, b);
            wide_load_rows<NT, K>(A, g, L, R, inv);
            wide_load_flux<K>(A, g, L, x);
            wide_spmv<NT, K>(A, R, L, x, t, S);
#pragma unroll
            for (int k = 0; k < K; k++) sb[L.c[k]] = b[k];  // 0 at dummies
            // ||A x - b||^2 of the stored flux: BiCGSTAB's initial residual
            double rg;
            n_bicg += wide_bicgstab<NT, K>(A, R, L, inv, x, t, b, tol, 150, S, sb, rg);
            ri += rg;
#pragma unroll
            for (int k = 0; k < K; k++)
                if (L.on[k]) A.cflux[(size_t)g * A.nc + L.c[k]] = x[k];
        }
        ri = sqrt(ri) / (A.nc * G);
        // the fission total and the source change in one sum (each value
        // summed as alone)
        const double in[2] = {lane_fission_total<K>(A, L), lane_source_change<K>(L, fsc, fs_old)};
        double out[2];
        wide_sum<NT, 2>(in, S, out);
        tfis = out[0];
        const double k_old = k_eff;
        k_eff = k_eff * tfis / tfis_old;
        const double psi_err = sqrt(out[1]);
        if (((fabs(k_eff - k_old) < A.k_tol) && (psi_err < A.psi_tol) && (ri / r0 < A.rr)) || (iter > A.max_iter))
            break;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
        if (L.on[k])
            for (int g = 0; g < G; g++) {
                double &f = A.cflux[(size_t)g * A.nc + L.c[k]];
                if (f < 0.0) f = -f;
            }
    if (threadIdx.x == 0) {
        A.k[0] = k_eff;
        A.info[0] = iter;
        A.info[2] = n_bicg;
    }
}

template <int NT, int K> __global__ __launch_bounds__(NT) void k_cmfd_solve_wide(SolveArgs A)
{
    extern __shared__ double lds[];
    __shared__ double red[2 * wide_red_bank<NT>()];
    cmfd_solve_wide<NT, K>(A, lds, red);
}

// Launch the wide solve at k cells per lane (K <= k <= KMAX).
template <int NT, int K, int KMAX> void launch_wide(int k, const SolveArgs &a, size_t lds, hipStream_t s)
{
    if (k == K)
        hipLaunchKernelGGL((k_cmfd_solve_wide<NT, K>), dim3(1), dim3(NT), lds, s, a);
    else if constexpr (K < KMAX)
        launch_wide<NT, K + 1, KMAX>(k, a, lds, s);
}

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

template <typename T> void dalloc(T *&p, size_t n)
{
    CMFD_CHECK(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)));
}

}  // namespace

CmfdSolver::~CmfdSolver() { free_all(); }

void CmfdSolver::free_all()
{
    double **ds[] = {&d_dx_, &d_dy_, &d_cvol_, &d_cflux_, &d_htr_, &d_hnf_, &d_hf_, &d_hch_, &d_hrm_,
                     &d_hsc_, &d_dtil_, &d_stil_, &d_dhat_, &d_shat_, &d_coef_, &d_gscr_, &d_part_,
                     &d_part_old_};
    for (double **p : ds) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (d_info_) (void)hipFree(d_info_);
    d_info_ = nullptr;
    if (d_smask_) (void)hipFree(d_smask_);
    d_smask_ = nullptr;
}

void CmfdSolver::update_xs(const mocgpu_problem &p, hipStream_t s)
{
    if (!d_smask_) return;
    if (p.n_group != G_) throw std::runtime_error("CMFD: cross sections of another group count");
    std::vector<uint64_t> m(G_, 0);
    for (int mat = 0; mat < p.n_mat; mat++)
        for (int g = 0; g < G_; g++)
            for (int gg = 0; gg < G_; gg++)
                if (gg != g && p.xs_sc[((size_t)mat * G_ + g) * G_ + gg] != 0.0) m[g] |= 1ull << gg;
    CMFD_CHECK(hipMemcpyAsync(d_smask_, m.data(), m.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    CMFD_CHECK(hipStreamSynchronize(s));
}

void CmfdSolver::setup(const mocgpu_problem &p, const mocgpu_options &opt, hipStream_t s)
{
    free_all();
    opt_ = opt;
    nx_ = p.nx;
    ny_ = p.ny;
    nc_ = p.nx * p.ny;
    G_ = p.n_group;
    hz_ = p.hz;
    n_surf_ = (nx_ + 1) * ny_ + (ny_ + 1) * nx_ + 2 * nx_ * ny_;
    for (int i = 0; i < 6; i++) bc_[i] = p.bc[i];
    if (!opt.use_cmfd) return;
    if (opt.cmfd_solver != 0) throw std::runtime_error("only the BiCGSTAB CMFD solver is supported");
    if (nc_ > kMaxCells)
        throw std::runtime_error("CMFD mesh of " + std::to_string(nc_) + " cells exceeds the supported " +
                                 std::to_string(kMaxCells));
    for (int i = 0; i < 6; i++)
        if (bc_[i] != MOCGPU_BC_VACUUM && bc_[i] != MOCGPU_BC_REFLECT)
            throw std::runtime_error("CMFD supports vacuum and reflective boundaries only");
    {
        // the solve keeps its operand banks (and rows or vectors) in LDS
        int dev = 0, lds_max = 0;
        CMFD_CHECK(hipGetDevice(&dev));
        CMFD_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t need = cmfd_wide(nc_) ? cmfd_wide_lds_bytes(nc_, nx_) + 2 * 32 * sizeof(double)
                                           : cmfd_lds_bytes(nc_) + 2 * red_bank<kCmfdThreads>() * sizeof(double);
        if (need > (size_t)lds_max)
            throw std::runtime_error("CMFD mesh of " + std::to_string(nc_) + " cells needs " + std::to_string(need) +
                                     " bytes of LDS per workgroup; the device allows " + std::to_string(lds_max));
    }
    std::vector<double> dx(nx_), dy(ny_), cv(nc_);
    for (int i = 0; i < nx_; i++) dx[i] = p.x_vec[i + 1] - p.x_vec[i];
    for (int j = 0; j < ny_; j++) dy[j] = p.y_vec[j + 1] - p.y_vec[j];
    for (int c = 0; c < nc_; c++) cv[c] = dx[c % nx_] * dy[c / nx_] * hz_;
    dalloc(d_dx_, nx_);
    dalloc(d_dy_, ny_);
    dalloc(d_cvol_, nc_);
    CMFD_CHECK(hipMemcpyAsync(d_dx_, dx.data(), nx_ * 8, hipMemcpyHostToDevice, s));
    CMFD_CHECK(hipMemcpyAsync(d_dy_, dy.data(), ny_ * 8, hipMemcpyHostToDevice, s));
    CMFD_CHECK(hipMemcpyAsync(d_cvol_, cv.data(), nc_ * 8, hipMemcpyHostToDevice, s));
    const size_t ncg = (size_t)nc_ * G_;
    dalloc(d_cflux_, ncg);
    dalloc(d_htr_, ncg);
    dalloc(d_hnf_, ncg);
    dalloc(d_hf_, ncg);
    dalloc(d_hch_, ncg);
    dalloc(d_hrm_, ncg);
    dalloc(d_hsc_, ncg * G_);
    dalloc(d_dtil_, (size_t)n_surf_ * G_);
    dalloc(d_stil_, (size_t)n_surf_ * G_);
    dalloc(d_dhat_, (size_t)n_surf_ * G_);
    dalloc(d_shat_, (size_t)n_surf_ * G_);
    dalloc(d_coef_, 5 * ncg);
    // [3][nc] for the solve of batches, [2][K * lanes] for the wide solve
    dalloc(d_gscr_, std::max(3 * (size_t)nc_, 2 * (size_t)cmfd_wide_k(nc_) * cmfd_wide_lanes(nc_)));
    dalloc(d_part_, 2 * (size_t)n_surf_ * G_);
    dalloc(d_part_old_, 2 * (size_t)n_surf_ * G_);
    dalloc(d_info_, 3);
    if (G_ > 64) throw std::runtime_error("CMFD: at most 64 groups");
    dalloc(d_smask_, G_);
    update_xs(p, s);
    initialize(s);
}

void CmfdSolver::initialize(hipStream_t s)
{
    n_solve_ = 0;
    has_old_partial_ = false;
    if (!d_info_) return;
    hipLaunchKernelGGL(k_zero_info, dim3(1), dim3(64), 0, s, d_info_);
    const size_t np = (size_t)n_surf_ * G_ * 2;
    CMFD_CHECK(hipMemsetAsync(d_part_, 0, np * sizeof(double), s));
    CMFD_CHECK(hipMemsetAsync(d_part_old_, 0, np * sizeof(double), s));
}


// The arguments of one problem's CMFD kernels (shared by enqueue_step and the
// batch launches over many problems).
struct CmfdAccess {
    static CoupleArgs couple(const CmfdSolver &c, const CmfdView &v)
    {
        CoupleArgs ca{};
        ca.m = Grid{c.nx_, c.ny_, c.hz_, c.d_dx_, c.d_dy_};
        ca.G = c.G_;
        ca.n_surf = c.n_surf_;
        ca.has_radial = v.has_radial;
        ca.radial_flag = nullptr;
        ca.bc_lo[0] = c.bc_[WEST];
        ca.bc_hi[0] = c.bc_[EAST];
        ca.bc_lo[1] = c.bc_[SOUTH];
        ca.bc_hi[1] = c.bc_[NORTH];
        ca.bc_lo[2] = c.bc_[BOTTOM];
        ca.bc_hi[2] = c.bc_[TOP];
        ca.htr = c.d_htr_;
        ca.cflux = c.d_cflux_;
        ca.cur = v.cur;
        ca.sfl = v.sfl;
        ca.dtil = c.d_dtil_;
        ca.stil = c.d_stil_;
        ca.dhat = c.d_dhat_;
        ca.shat = c.d_shat_;
        ca.part = c.d_part_;
        ca.part_old = c.d_part_old_;
        return ca;
    }
    static SolveArgs solve(const CmfdSolver &c, const CmfdView &v)
    {
        SolveArgs a{};
        a.nc = c.nc_;
        a.nx = c.nx_;
        a.G = c.G_;
        a.max_iter = c.opt_.cmfd_max_iter;
        a.k_tol = c.opt_.cmfd_k_tol;
        a.psi_tol = c.opt_.cmfd_psi_tol;
        a.rr = c.opt_.cmfd_resid_reduction;
        a.coef = c.d_coef_;
        a.hnf = c.d_hnf_;
        a.hch = c.d_hch_;
        a.hsc = c.d_hsc_;
        a.cvol = c.d_cvol_;
        a.smask = c.d_smask_;
        a.cflux = c.d_cflux_;
        a.gscr = c.d_gscr_;
        a.k = v.k;
        a.info = c.d_info_;
        return a;
    }
    static HomogArrays homog(const CmfdSolver &c)
    {
        return HomogArrays{c.d_htr_, c.d_hnf_, c.d_hf_, c.d_hch_, c.d_hrm_, c.d_hsc_};
    }
    static int n_cell(const CmfdSolver &c) { return c.nc_; }
    static int n_surf(const CmfdSolver &c) { return c.n_surf_; }
    static double *hrm(const CmfdSolver &c) { return c.d_hrm_; }
    static double *coef(const CmfdSolver &c) { return c.d_coef_; }
    static bool enabled(const CmfdSolver &c) { return c.d_cflux_ != nullptr; }
};

void CmfdSolver::homogenize(const CmfdView &v, hipStream_t s)
{
    if (!d_cflux_) throw std::runtime_error("CMFD is not enabled for this handle (use_cmfd = 0)");
    const int npg = v.npin * v.G;
    hipLaunchKernelGGL(k_homogenize, dim3(nblk(npg)), dim3(256), 0, s, v, CmfdAccess::homog(*this));
    CMFD_CHECK(hipGetLastError());
}

void CmfdSolver::enqueue_step(const CmfdView &v, hipStream_t s)
{
    if (!d_cflux_) throw std::runtime_error("CMFD is not enabled for this handle (use_cmfd = 0)");
    const int npg = v.npin * v.G;
    hipLaunchKernelGGL(k_zero_info, dim3(1), dim3(64), 0, s, d_info_);
    hipLaunchKernelGGL(k_pin_flux_all, dim3(nblk(npg)), dim3(256), 0, s, v, d_cflux_);
    hipLaunchKernelGGL(k_homogenize, dim3(nblk(npg)), dim3(256), 0, s, v, CmfdAccess::homog(*this));
    const CoupleArgs ca = CmfdAccess::couple(*this, v);
    hipLaunchKernelGGL(k_couple, dim3(nblk((size_t)n_surf_ * G_)), dim3(256), 0, s, ca);
    hipLaunchKernelGGL(k_assemble, dim3(nblk((size_t)nc_ * G_)), dim3(256), 0, s, ca.m, G_, d_hrm_, d_dtil_,
                       d_dhat_, d_coef_);
    const SolveArgs a = CmfdAccess::solve(*this, v);
    if (cmfd_wide(nc_)) {
        const size_t lds = cmfd_wide_lds_bytes(nc_, nx_);
        if (cmfd_wide_lanes(nc_) == 512)
            launch_wide<512, 3, 6>(cmfd_wide_k(nc_), a, lds, s);
        else
            launch_wide<1024, 4, 6>(cmfd_wide_k(nc_), a, lds, s);
    } else {
        const size_t lds = cmfd_lds_bytes(nc_);  // operands, fs_old, BiCGSTAB vectors
        switch (cmfd_cells_per_lane(nc_)) {
        case 1:
            switch (cmfd_threads(nc_)) {
            case 64: hipLaunchKernelGGL((k_cmfd_solve<1, 64>), dim3(1), dim3(64), lds, s, a); break;
            case 128: hipLaunchKernelGGL((k_cmfd_solve<1, 128>), dim3(1), dim3(128), lds, s, a); break;
            case 192: hipLaunchKernelGGL((k_cmfd_solve<1, 192>), dim3(1), dim3(192), lds, s, a); break;
            case 256: hipLaunchKernelGGL((k_cmfd_solve<1, 256>), dim3(1), dim3(256), lds, s, a); break;
            case 320: hipLaunchKernelGGL((k_cmfd_solve<1, 320>), dim3(1), dim3(320), lds, s, a); break;
            case 384: hipLaunchKernelGGL((k_cmfd_solve<1, 384>), dim3(1), dim3(384), lds, s, a); break;
            case 448: hipLaunchKernelGGL((k_cmfd_solve<1, 448>), dim3(1), dim3(448), lds, s, a); break;
            default: hipLaunchKernelGGL((k_cmfd_solve<1, 512>), dim3(1), dim3(512), lds, s, a); break;
            }
            break;
        case 2: hipLaunchKernelGGL((k_cmfd_solve<2, kCmfdThreads>), dim3(1), dim3(kCmfdThreads), lds, s, a); break;
        default: throw std::runtime_error("the 512-lane coarse solve takes at most 1024 cells");
        }
    }
    hipLaunchKernelGGL(k_store_currents, dim3(nblk((size_t)n_surf_ * G_)), dim3(256), 0, s, ca, v.cur, v.sfl);
    has_old_partial_ = n_solve_ > 0;  // cmfd.cpp:566, before n_solve_++ (:274)
    hipLaunchKernelGGL(k_set_pin_flux_all, dim3(nblk(npg)), dim3(256), 0, s, v, d_cflux_, d_info_ + 1);
    CMFD_CHECK(hipGetLastError());
    n_solve_++;
}

void CmfdSolver::read_info(hipStream_t s, int *iterations, int *bad_flux, int *bicgstab_iterations)
{
    int info[3] = {0, 0, 0};
    if (d_info_) {
        CMFD_CHECK(hipMemcpyAsync(info, d_info_, sizeof(info), hipMemcpyDeviceToHost, s));
        CMFD_CHECK(hipStreamSynchronize(s));
    }
    if (iterations) *iterations = info[0];
    if (bad_flux) *bad_flux = info[1];
    if (bicgstab_iterations) *bicgstab_iterations = info[2];
}

void CmfdSolver::matrix_csr(int g, hipStream_t s, std::vector<int32_t> &row_ptr, std::vector<int32_t> &col,
                            std::vector<double> &val)
{
    std::vector<double> coef((size_t)5 * nc_ * G_);
    CMFD_CHECK(hipMemcpyAsync(coef.data(), d_coef_, coef.size() * 8, hipMemcpyDeviceToHost, s));
    CMFD_CHECK(hipStreamSynchronize(s));
    const size_t pl = (size_t)G_ * nc_;
    row_ptr.assign(1, 0);
    col.clear();
    val.clear();
    for (int c = 0; c < nc_; c++) {
        const int ix = c % nx_, iy = c / nx_;
        const size_t i = (size_t)g * nc_ + c;
        const int nb[5] = {iy > 0 ? c - nx_ : -1, ix > 0 ? c - 1 : -1, c, ix < nx_ - 1 ? c + 1 : -1,
                           iy < ny_ - 1 ? c + nx_ : -1};
        for (int k = 0; k < 5; k++)
            if (nb[k] >= 0) {
                col.push_back(nb[k]);
                val.push_back(coef[k * pl + i]);
            }
        row_ptr.push_back((int32_t)col.size());
    }
}

void CmfdSolver::coarse_flux(hipStream_t s, std::vector<double> &out)
{
    // the device keeps [G][nc]; the caller gets [nc][G]
    std::vector<double> t((size_t)nc_ * G_, 0.0);
    CMFD_CHECK(hipMemcpyAsync(t.data(), d_cflux_, t.size() * 8, hipMemcpyDeviceToHost, s));
    CMFD_CHECK(hipStreamSynchronize(s));
    out.assign(t.size(), 0.0);
    for (int c = 0; c < nc_; c++)
        for (int g = 0; g < G_; g++) out[(size_t)c * G_ + g] = t[(size_t)g * nc_ + c];
}

// --- the CMFD step of many problems (batch engine) ------------------------------

namespace {

struct CmfdCase {
    int32_t bc;             // index of the problem in the batch status array
    int32_t npg, nsg, ncg;  // elements: pins x G, surfaces x G, cells x G
    int32_t wide_k;         // cells per lane of its wide solve (mesh over 1024 cells), else 0
    CmfdView v;
    CoupleArgs ca;
    SolveArgs a;
    HomogArrays h;
    double *hrm, *coef;
    int *info;
};

struct CaseBlock {
    int32_t c, first;
};

__device__ __forceinline__ bool case_done(const CmfdCase &d, const int32_t *status)
{
    return status[2 * d.bc + 1] != 0;
}

__global__ void k_cmfd_b_zero(const CmfdCase *__restrict__ cs, int n, const int32_t *__restrict__ status)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || case_done(cs[i], status)) return;
    for (int q = 0; q < 3; q++) cs[i].info[q] = 0;
}

// stage 0: pin flux, 1: homogenize, 5: rescale (pins x G);
// 2: couple, 4: store currents (surfaces x G); 3: assemble (cells x G)
template <int STAGE>
__global__ void k_cmfd_b_elem(const CmfdCase *__restrict__ cs, const CaseBlock *__restrict__ bm,
                              const int32_t *__restrict__ status)
{
    const CaseBlock b = bm[blockIdx.x];
    const int c = __builtin_amdgcn_readfirstlane(b.c);
    const CmfdCase &d = cs[c];
    if (case_done(d, status)) return;
    const int t = __builtin_amdgcn_readfirstlane(b.first) + threadIdx.x;
    if (STAGE == 0) pin_flux_item(d.v, d.a.cflux, t);
    if (STAGE == 1) homogenize_item(d.v, d.h, t);
    if (STAGE == 2) couple_item(d.ca, t);
    if (STAGE == 3) assemble_item(d.ca.m, d.ca.G, d.hrm, d.ca.dtil, d.ca.dhat, d.coef, t);
    if (STAGE == 4) store_currents_item(d.ca, d.v.cur, d.v.sfl, t);
    if (STAGE == 5) set_pin_flux_item(d.v, d.a.cflux, d.info + 1, t);
}

template <int K, int NT, bool RL = false>
__global__ __launch_bounds__(NT) void k_cmfd_b_solve(const CmfdCase *__restrict__ cs, const int32_t *__restrict__ status)
{
    extern __shared__ double lds[];
    __shared__ double red[2 * red_bank<NT>()];
    const CmfdCase &d = cs[blockIdx.x];
    if (case_done(d, status)) return;
    cmfd_solve<K, NT, RL>(d.a, lds, red);
}

// A batch that holds meshes of more than 1024 cells: each of those runs the
// wide solve at its own cells per lane (512 lanes, as a single problem of that
// size does), the others the 512-lane solve at K cells per lane (the largest
// of them, up to 1024 cells). Which solve a problem gets depends only on its
// own mesh, and neither solve's arithmetic depends on K or on the lane count
// beyond the problem's own, so a problem's result does not depend on the
// others in the batch.
template <int K>
__global__ __launch_bounds__(kCmfdThreads) void k_cmfd_b_solve_mixed(const CmfdCase *__restrict__ cs,
                                                                     const int32_t *__restrict__ status)
{
    static_assert(wide_red_bank<kCmfdThreads>() >= red_bank<kCmfdThreads>(), "one reduction array serves both solves");
    extern __shared__ double lds[];
    __shared__ double red[2 * wide_red_bank<kCmfdThreads>()];
    const CmfdCase &d = cs[blockIdx.x];
    if (case_done(d, status)) return;
    switch (__builtin_amdgcn_readfirstlane(d.wide_k)) {
    case 0: cmfd_solve<K, kCmfdThreads>(d.a, lds, red); break;
    case 3: cmfd_solve_wide<kCmfdThreads, 3>(d.a, lds, red); break;
    case 4: cmfd_solve_wide<kCmfdThreads, 4>(d.a, lds, red); break;
    case 5: cmfd_solve_wide<kCmfdThreads, 5>(d.a, lds, red); break;
    default: cmfd_solve_wide<kCmfdThreads, 6>(d.a, lds, red); break;
    }
}

std::vector<CaseBlock> case_blocks(const std::vector<int> &counts)
{
    std::vector<CaseBlock> m;
    for (size_t c = 0; c < counts.size(); c++)
        for (int first = 0; first < counts[c]; first += 256) m.push_back(CaseBlock{(int32_t)c, first});
    return m;
}

}  // namespace

struct CmfdBatch::Impl {
    CmfdCase *cases = nullptr;
    CaseBlock *bm[3] = {nullptr, nullptr, nullptr};  // pins, surfaces, cells
    int nbm[3] = {0, 0, 0};
    int kmax = 1;          // cells per lane of the largest mesh (of the narrow ones when mixed)
    bool mixed = false;    // meshes of more than 1024 cells present: k_cmfd_b_solve_mixed
    int nt = kCmfdThreads; // lanes of the solve workgroups
    size_t lds = 0;
    void release()
    {
        if (cases) (void)hipFree(cases);
        for (auto &p : bm)
            if (p) (void)hipFree(p);
        cases = nullptr;
        for (auto &p : bm) p = nullptr;
    }
    ~Impl() { release(); }
};

CmfdBatch::CmfdBatch() : impl_(new Impl) {}
CmfdBatch::~CmfdBatch() { delete impl_; }

void CmfdBatch::build(const std::vector<CmfdBatchEntry> &entries, hipStream_t s)
{
    CMFD_CHECK(hipStreamSynchronize(s));
    impl_->release();
    n_ = 0;
    if (entries.empty()) return;
    std::vector<CmfdCase> host;
    std::vector<int> np, ns, nc;
    int ncmax = 0, ncmax_narrow = 0;
    size_t lds_wide = 0;
    for (const CmfdBatchEntry &e : entries) {
        const CmfdSolver &sv = *e.solver;
        if (!CmfdAccess::enabled(sv)) throw std::runtime_error("CMFD is not enabled for a batch problem");
        CmfdCase d{};
        d.bc = e.batch_index;
        d.v = e.view;
        d.ca = CmfdAccess::couple(sv, e.view);
        d.ca.radial_flag = e.radial_flag;
        d.a = CmfdAccess::solve(sv, e.view);
        d.h = CmfdAccess::homog(sv);
        d.hrm = CmfdAccess::hrm(sv);
        d.coef = CmfdAccess::coef(sv);
        d.info = d.a.info;
        d.npg = e.view.npin * e.view.G;
        d.nsg = CmfdAccess::n_surf(sv) * e.view.G;
        d.ncg = CmfdAccess::n_cell(sv) * e.view.G;
        const int ncell = CmfdAccess::n_cell(sv);
        d.wide_k = 0;
        if (cmfd_wide(ncell)) {
            d.wide_k = cmfd_wide_k(ncell);
            lds_wide = std::max(lds_wide, cmfd_wide_lds_bytes(ncell, d.a.nx));
        } else {
            ncmax_narrow = std::max(ncmax_narrow, ncell);
        }
        host.push_back(d);
        np.push_back(d.npg);
        ns.push_back(d.nsg);
        nc.push_back(d.ncg);
        ncmax = std::max(ncmax, ncell);
    }
    n_ = (int)host.size();
    dalloc(impl_->cases, host.size());
    CMFD_CHECK(hipMemcpyAsync(impl_->cases, host.data(), host.size() * sizeof(CmfdCase), hipMemcpyHostToDevice, s));
    const std::vector<int> *cnt[3] = {&np, &ns, &nc};
    std::vector<CaseBlock> maps[3];
    for (int k = 0; k < 3; k++) {
        maps[k] = case_blocks(*cnt[k]);
        impl_->nbm[k] = (int)maps[k].size();
        dalloc(impl_->bm[k], maps[k].size());
        CMFD_CHECK(hipMemcpyAsync(impl_->bm[k], maps[k].data(), maps[k].size() * sizeof(CaseBlock),
                                  hipMemcpyHostToDevice, s));
    }
    if (ncmax > CmfdSolver::kMaxBatchCells)
        throw std::runtime_error("a batched CMFD problem holds at most " + std::to_string(CmfdSolver::kMaxBatchCells) +
                                 " coarse cells; this batch has one of " + std::to_string(ncmax));
    impl_->mixed = lds_wide > 0;
    if (impl_->mixed) {
        // the narrow meshes (up to 1024 cells: 1 or 2 cells per lane) on
        // kCmfdThreads lanes, whose LDS vectors are K * kCmfdThreads long
        impl_->kmax = cmfd_cells_per_lane(std::max(1, ncmax_narrow));
        impl_->nt = kCmfdThreads;
        impl_->lds = std::max(lds_wide, (size_t)kCmfdLdsVectors * impl_->kmax * kCmfdThreads * sizeof(double));
    } else {
        impl_->kmax = cmfd_cells_per_lane(ncmax);
        impl_->nt = cmfd_threads(ncmax);
        impl_->lds = cmfd_lds_bytes(ncmax);
    }
    CMFD_CHECK(hipStreamSynchronize(s));
}

void CmfdBatch::enqueue(const int32_t *status, hipStream_t s) const
{
    if (n_ == 0) return;
    const Impl &m = *impl_;
    hipLaunchKernelGGL(k_cmfd_b_zero, dim3(nblk(n_, 64)), dim3(64), 0, s, m.cases, n_, status);
    hipLaunchKernelGGL(k_cmfd_b_elem<0>, dim3(m.nbm[0]), dim3(256), 0, s, m.cases, m.bm[0], status);
    hipLaunchKernelGGL(k_cmfd_b_elem<1>, dim3(m.nbm[0]), dim3(256), 0, s, m.cases, m.bm[0], status);
    hipLaunchKernelGGL(k_cmfd_b_elem<2>, dim3(m.nbm[1]), dim3(256), 0, s, m.cases, m.bm[1], status);
    hipLaunchKernelGGL(k_cmfd_b_elem<3>, dim3(m.nbm[2]), dim3(256), 0, s, m.cases, m.bm[2], status);
    const dim3 grid(n_), block(kCmfdThreads);
    if (m.mixed) {
        if (m.kmax == 1)
            hipLaunchKernelGGL(k_cmfd_b_solve_mixed<1>, grid, block, m.lds, s, m.cases, status);
        else
            hipLaunchKernelGGL(k_cmfd_b_solve_mixed<2>, grid, block, m.lds, s, m.cases, status);
    } else {
        switch (m.kmax) {
        case 1:
            switch (m.nt) {
            case 64: hipLaunchKernelGGL((k_cmfd_b_solve<1, 64>), grid, dim3(64), m.lds, s, m.cases, status); break;
            case 128: hipLaunchKernelGGL((k_cmfd_b_solve<1, 128>), grid, dim3(128), m.lds, s, m.cases, status); break;
            case 192: hipLaunchKernelGGL((k_cmfd_b_solve<1, 192>), grid, dim3(192), m.lds, s, m.cases, status); break;
            case 256: hipLaunchKernelGGL((k_cmfd_b_solve<1, 256>), grid, dim3(256), m.lds, s, m.cases, status); break;
            case 320: hipLaunchKernelGGL((k_cmfd_b_solve<1, 320>), grid, dim3(320), m.lds, s, m.cases, status); break;
            case 384: hipLaunchKernelGGL((k_cmfd_b_solve<1, 384>), grid, dim3(384), m.lds, s, m.cases, status); break;
            case 448: hipLaunchKernelGGL((k_cmfd_b_solve<1, 448>), grid, dim3(448), m.lds, s, m.cases, status); break;
            default: hipLaunchKernelGGL((k_cmfd_b_solve<1, 512>), grid, dim3(512), m.lds, s, m.cases, status); break;
            }
            break;
        case 2: hipLaunchKernelGGL((k_cmfd_b_solve<2, kCmfdThreads>), grid, block, m.lds, s, m.cases, status); break;
        default: throw std::runtime_error("the 512-lane coarse solve takes at most 1024 cells");
        }
    }
    hipLaunchKernelGGL(k_cmfd_b_elem<4>, dim3(m.nbm[1]), dim3(256), 0, s, m.cases, m.bm[1], status);
    hipLaunchKernelGGL(k_cmfd_b_elem<5>, dim3(m.nbm[0]), dim3(256), 0, s, m.cases, m.bm[0], status);
}

}  // namespace gpu
}  // namespace mocc_amd
