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
//     test -- runs in ONE workgroup of kCmfdThreads lanes. Each lane owns
//     K cells (c = lane + j * kCmfdThreads) and keeps their BiCGSTAB vectors
//     in registers; a matrix-vector product stages the operand in LDS, and
//     dot products are wave butterflies plus a fixed-order sum over the
//     waves, so every lane sees the same scalars and the iteration is
//     deterministic. No host round trip happens inside a CMFD solve.
#include "cmfd.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace mocc_amd {
namespace gpu {

namespace {

constexpr int kCmfdThreads = 512;
constexpr int kCmfdWaves = kCmfdThreads / 64;
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

// MoCSweeper::get_pin_flux_1g (PIN_PLANE, moc_sweeper.cpp:293-326) for all groups.
__global__ void k_pin_flux_all(CmfdView v, double *cflux)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= v.npin * v.G) return;
    const int ip = t / v.G, g = t % v.G;
    double vs = 0.0, pf = 0.0;
    for (int r = v.pin_off[ip]; r < v.pin_off[ip + 1]; r++) {
        vs += v.vol[r];
        pf += v.phi[(size_t)r * v.G + g] * v.vol[r];
    }
    cflux[(size_t)v.pin_cell[ip] * v.G + g] = pf / vs;
}

// XSMeshHomogenized::homogenize_region_flux: flux-area weighted Sigma_t,
// nu Sigma_f, Sigma_f; chi weighted by the fission source; Sigma_s[g <- g']
// weighted by phi_g'. One thread per (pin, group).
__global__ void k_homogenize(CmfdView v, double *htr, double *hnf, double *hf, double *hch, double *hrm,
                             double *hsc)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= v.npin * v.G) return;
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
    hnf[(size_t)cell * G + g] = nf / fvs;
    hf[(size_t)cell * G + g] = fi / fvs;
    hch[(size_t)cell * G + g] = (fs_sum > 0.0) ? ch / fs_sum : ch;
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
        hsc[((size_t)cell * G + g) * G + gg] = sc;
        if (gg == g) self = sc;
    }
    hrm[(size_t)cell * G + g] = htr_g - self;
}

struct CoupleArgs {
    Grid m;
    int G, n_surf, has_radial;
    int bc_lo[3], bc_hi[3];  // [normal]: low-side and high-side boundary kinds
    const double *htr, *cflux, *cur, *sfl;
    double *dtil, *stil, *dhat, *shat;
};

// CMFD::setup_solve, surface couplings (cmfd.cpp:423-490)
__global__ void k_couple(CoupleArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_surf * a.G) return;
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
    if (norm != 2 && a.has_radial) {
        const double j = a.cur[(size_t)is * G + g];
        const double sf = a.sfl[(size_t)is * G + g];
        const double fl = c1 >= 0 ? a.cflux[(size_t)c1 * G + g] : 0.0;
        const double fr = c2 >= 0 ? a.cflux[(size_t)c2 * G + g] : 0.0;
        dh = (j + dt * (fr - fl)) / (fl + fr);
        if (!isfinite(dh)) dh = 0.0;
        sh = (c1 >= 0) ? (sf - st * fl - (1.0 - st) * fr) / (fl + fr) : (sf - st * fr) / fr;
    }
    a.dhat[(size_t)is * G + g] = dh;
    a.shat[(size_t)is * G + g] = sh;
}

// CMFD::setup_solve, matrix rows (cmfd.cpp:506-544): one thread per (cell, g)
__global__ void k_assemble(Grid m, int G, const double *hrm, const double *dtil, const double *dhat, double *coef)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
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

// CMFD::store_currents (cmfd.cpp:555-638): the net current and surface flux
// the coarse solution implies.
__global__ void k_store_currents(CoupleArgs a, double *cur, double *sfl)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_surf * a.G) return;
    const int is = t / a.G, g = t % a.G, G = a.G;
    int c1, c2;
    neigh_cells(a.m, is, c1, c2);
    const double fr = c2 >= 0 ? a.cflux[(size_t)c2 * G + g] : 0.0;
    const double fl = c1 >= 0 ? a.cflux[(size_t)c1 * G + g] : 0.0;
    const double dh = a.dhat[(size_t)is * G + g], dt = a.dtil[(size_t)is * G + g];
    const double j = -dt * (fr - fl) + dh * (fr + fl);
    const double sh = a.shat[(size_t)is * G + g], st = a.stil[(size_t)is * G + g];
    const double sf = (c1 >= 0) ? st * fl + (1.0 - st) * fr + sh * (fl + fr) : st * fr + sh * (fl + fr);
    sfl[(size_t)is * G + g] = sf;
    cur[(size_t)is * G + g] = j;
}

// MoCSweeper::set_pin_flux_1g (PIN_PLANE, moc_sweeper.cpp:360-433) for all
// groups; a non-positive coarse flux raises the flag (the reference throws).
__global__ void k_set_pin_flux_all(CmfdView v, const double *cflux, int *bad)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= v.npin * v.G) return;
    const int ip = t / v.G, g = t % v.G;
    const double pn = cflux[(size_t)v.pin_cell[ip] * v.G + g];
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
    double *cflux;  // [nc][G]
    double *k;
    int *info;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum over the workgroup; every lane returns the same value (lane 0 of each
// wave publishes, all lanes add the wave sums in wave order).
__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kCmfdWaves; w++) s += red[w];
    return s;
}

__device__ __forceinline__ void block_sum2(double a, double b, double *red, double &sa, double &sb)
{
    a = wave_sum(a);
    b = wave_sum(b);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = a;
        red[kCmfdWaves + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    sa = 0.0;
    sb = 0.0;
#pragma unroll
    for (int w = 0; w < kCmfdWaves; w++) {
        sa += red[w];
        sb += red[kCmfdWaves + w];
    }
}

template <int K> struct Lane {
    int c[K];
    bool on[K];
};

// v = A y for group g: y staged in LDS, rows summed in sorted column order
// (south, west, diagonal, east, north) as a CSR product would. y and v may be
// the same array; y stays readable in buf until the next product.
template <int K>
__device__ __forceinline__ void spmv(const SolveArgs &A, int g, const Lane<K> &L, const double (&y)[K],
                                     double (&v)[K], double *buf)
{
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++)
        if (L.on[k]) buf[L.c[k]] = y[k];
    __syncthreads();
    const size_t pl = (size_t)A.G * A.nc;
    // reload the coefficients from cache at every product rather than keep
    // 5K of them live in registers across the iteration
    const double *cf = A.coef;
    __asm__ volatile("" : "+s"(cf));
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        if (L.on[k]) {
            const int c = L.c[k];
            const int ix = c % A.nx;
            const size_t i = (size_t)g * A.nc + c;
            if (c >= A.nx) s += cf[0 * pl + i] * buf[c - A.nx];
            if (ix > 0) s += cf[1 * pl + i] * buf[c - 1];
            s += cf[2 * pl + i] * buf[c];
            if (ix < A.nx - 1) s += cf[3 * pl + i] * buf[c + 1];
            if (c + A.nx < A.nc) s += cf[4 * pl + i] * buf[c + A.nx];
        }
        v[k] = s;
    }
}

// Jacobi-preconditioned BiCGSTAB with the relative test |r| <= tol |b| and a
// restart when rho vanishes (van der Vorst 1992, the iteration of Eigen's
// BiCGSTAB<SparseMatrix<double>> that cmfd.cpp:291 calls). The preconditioned
// directions y = M^-1 p and z = M^-1 s are read back from the product's LDS
// operand, so x is updated as (x + alpha y) + w z. The loop-invariant vectors
// (the inverse diagonal, the shadow residual r0 and b) live in LDS slots that
// only the lane owning the cell touches, which keeps the rest in registers.
template <int K>
__device__ int bicgstab(const SolveArgs &A, int g, const Lane<K> &L, double (&x)[K], double tol, int max_iter,
                        double *buf, double *red, double *sinv, double *sr0, const double *sb)
{
    const size_t pl = (size_t)A.G * A.nc;
    double r[K], p[K], v[K], t[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (L.on[k]) {
            const double d = A.coef[2 * pl + (size_t)g * A.nc + L.c[k]];
            sinv[L.c[k]] = (d != 0.0) ? 1.0 / d : 1.0;
        }
    }
    spmv<K>(A, g, L, x, t, buf);
    double lr0 = 0.0, lb = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double bk = L.on[k] ? sb[L.c[k]] : 0.0;
        r[k] = bk - t[k];
        if (L.on[k]) sr0[L.c[k]] = r[k];
        p[k] = 0.0;
        v[k] = 0.0;
        lr0 += r[k] * r[k];
        lb += bk * bk;
    }
    double r0_sq, rhs_sq;
    block_sum2(lr0, lb, red, r0_sq, rhs_sq);
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
        double lrr = 0.0, lrho = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            lrr += r[k] * r[k];
            if (L.on[k]) lrho += sr0[L.c[k]] * r[k];
        }
        double rr, rho_new;
        block_sum2(lrr, lrho, red, rr, rho_new);
        if (!(rr > tol2 && i < max_iter)) break;
        const double rho_old = rho;
        rho = rho_new;
        if (fabs(rho) < eps2 * r0_sq) {
            spmv<K>(A, g, L, x, t, buf);
            double l = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                r[k] = (L.on[k] ? sb[L.c[k]] : 0.0) - t[k];
                if (L.on[k]) sr0[L.c[k]] = r[k];
                l += r[k] * r[k];
            }
            rho = r0_sq = block_sum(l, red);
            if (restarts++ == 0) i = 0;
        }
        const double beta = (rho / rho_old) * (alpha / w);
#pragma unroll
        for (int k = 0; k < K; k++) {
            p[k] = r[k] + beta * (p[k] - w * v[k]);
            t[k] = (L.on[k] ? sinv[L.c[k]] : 1.0) * p[k];  // y
        }
        spmv<K>(A, g, L, t, v, buf);
        double l = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (L.on[k]) l += sr0[L.c[k]] * v[k];
        alpha = rho / block_sum(l, red);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (L.on[k]) x[k] += alpha * buf[L.c[k]];
            r[k] = r[k] - alpha * v[k];                     // s
            t[k] = (L.on[k] ? sinv[L.c[k]] : 1.0) * r[k];  // z
        }
        spmv<K>(A, g, L, t, t, buf);  // t = A z
        double ltt = 0.0, lts = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            ltt += t[k] * t[k];
            lts += t[k] * r[k];
        }
        double tt, ts;
        block_sum2(ltt, lts, red, tt, ts);
        w = (tt > 0.0) ? ts / tt : 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (L.on[k]) x[k] += w * buf[L.c[k]];
            r[k] = r[k] - w * t[k];
        }
        ++i;
    }
    return i;
}

// CMFD source of group g: V (chi fs + sum_{g' != g} Sigma_s[g <- g'] phi_g')
// (cmfd.cpp:225-231), from this lane's own cells.
template <int K>
__device__ __forceinline__ void cmfd_source(const SolveArgs &A, int g, const Lane<K> &L, const double (&fsc)[K],
                                            double (&b)[K])
{
    const int G = A.G;
#pragma unroll
    for (int k = 0; k < K; k++) {
        b[k] = 0.0;
        if (!L.on[k]) continue;
        const int c = L.c[k];
        double s = A.hch[(size_t)c * G + g] * fsc[k];
        const double *row = A.hsc + ((size_t)c * G + g) * G;
        for (int gg = 0; gg < G; gg++) {
            if (gg == g) continue;
            s += row[gg] * A.cflux[(size_t)c * G + gg];
        }
        b[k] = s * A.cvol[c];
    }
}

template <int K>
__device__ __forceinline__ double fission_source(const SolveArgs &A, const Lane<K> &L, double k_eff, double (&fsc)[K])
{
    const double rk = 1.0 / k_eff;
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        if (L.on[k])
            for (int g = 0; g < A.G; g++) s += A.hnf[(size_t)L.c[k] * A.G + g] * A.cflux[(size_t)L.c[k] * A.G + g];
        tot += s;
        fsc[k] = s * rk;
    }
    return tot;
}

// ||A x - b||^2 of group g with x the stored coarse flux (cmfd.cpp:640-666)
template <int K>
__device__ __forceinline__ double residual2(const SolveArgs &A, int g, const Lane<K> &L, const double (&b)[K],
                                            double (&x)[K], double (&t)[K], double *buf)
{
#pragma unroll
    for (int k = 0; k < K; k++) x[k] = L.on[k] ? A.cflux[(size_t)L.c[k] * A.G + g] : 0.0;
    spmv<K>(A, g, L, x, t, buf);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) s += (t[k] - b[k]) * (t[k] - b[k]);
    return s;
}

// CMFD::solve (cmfd.cpp:171-279)
template <int K> __global__ __launch_bounds__(kCmfdThreads) void k_cmfd_solve(SolveArgs A)
{
    extern __shared__ double lds[];
    double *buf = lds;                // [nc] operand of the matrix-vector products
    double *fs_old = lds + A.nc;      // [nc] previous coarse fission source
    double *sinv = lds + 2 * A.nc;    // [nc] BiCGSTAB: inverse diagonal
    double *sr0 = lds + 3 * A.nc;     // [nc] BiCGSTAB: shadow residual
    double *sb = lds + 4 * A.nc;      // [nc] BiCGSTAB: right-hand side
    __shared__ double red[2 * kCmfdWaves];
    Lane<K> L;
#pragma unroll
    for (int k = 0; k < K; k++) {
        L.c[k] = threadIdx.x + k * kCmfdThreads;
        L.on[k] = L.c[k] < A.nc;
    }
    const int G = A.G;
    double k_eff = A.k[0];
    double fsc[K], b[K], x[K], t[K];
    double tfis = block_sum(fission_source<K>(A, L, k_eff, fsc), red);
    double r0 = 0.0;
    for (int g = 0; g < G; g++) {
        cmfd_source<K>(A, g, L, fsc, b);
        r0 += block_sum(residual2<K>(A, g, L, b, x, t, buf), red);
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
            ri += block_sum(residual2<K>(A, g, L, b, x, t, buf), red);
#pragma unroll
            for (int k = 0; k < K; k++)
                if (L.on[k]) sb[L.c[k]] = b[k];
            n_bicg += bicgstab<K>(A, g, L, x, tol, 150, buf, red, sinv, sr0, sb);
#pragma unroll
            for (int k = 0; k < K; k++)
                if (L.on[k]) A.cflux[(size_t)L.c[k] * G + g] = x[k];
        }
        ri = sqrt(ri) / (A.nc * G);
        double lf = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (L.on[k])
                for (int g = 0; g < G; g++) lf += A.hnf[(size_t)L.c[k] * G + g] * A.cflux[(size_t)L.c[k] * G + g];
        tfis = block_sum(lf, red);
        const double k_old = k_eff;
        k_eff = k_eff * tfis / tfis_old;
        double le = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (L.on[k]) le += (fsc[k] - fs_old[L.c[k]]) * (fsc[k] - fs_old[L.c[k]]);
        const double psi_err = sqrt(block_sum(le, red));
        if (((fabs(k_eff - k_old) < A.k_tol) && (psi_err < A.psi_tol) && (ri / r0 < A.rr)) || (iter > A.max_iter))
            break;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
        if (L.on[k])
            for (int g = 0; g < G; g++) {
                double &f = A.cflux[(size_t)L.c[k] * G + g];
                if (f < 0.0) f = -f;
            }
    if (threadIdx.x == 0) {
        A.k[0] = k_eff;
        A.info[0] = iter;
        A.info[2] = n_bicg;
    }
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
                     &d_hsc_, &d_dtil_, &d_stil_, &d_dhat_, &d_shat_, &d_coef_};
    for (double **p : ds) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (d_info_) (void)hipFree(d_info_);
    d_info_ = nullptr;
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
        // the solve keeps five cell vectors in LDS (k_cmfd_solve)
        int dev = 0, lds_max = 0;
        CMFD_CHECK(hipGetDevice(&dev));
        CMFD_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t need = 5 * (size_t)nc_ * sizeof(double) + 4 * kCmfdWaves * sizeof(double);
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
    dalloc(d_info_, 3);
    initialize(s);
}

void CmfdSolver::initialize(hipStream_t s)
{
    n_solve_ = 0;
    if (!d_info_) return;
    hipLaunchKernelGGL(k_zero_info, dim3(1), dim3(64), 0, s, d_info_);
}

void CmfdSolver::enqueue_step(const CmfdView &v, hipStream_t s)
{
    if (!d_cflux_) throw std::runtime_error("CMFD is not enabled for this handle (use_cmfd = 0)");
    const int npg = v.npin * v.G;
    hipLaunchKernelGGL(k_zero_info, dim3(1), dim3(64), 0, s, d_info_);
    hipLaunchKernelGGL(k_pin_flux_all, dim3(nblk(npg)), dim3(256), 0, s, v, d_cflux_);
    hipLaunchKernelGGL(k_homogenize, dim3(nblk(npg)), dim3(256), 0, s, v, d_htr_, d_hnf_, d_hf_, d_hch_, d_hrm_,
                       d_hsc_);
    CoupleArgs ca{};
    ca.m = Grid{nx_, ny_, hz_, d_dx_, d_dy_};
    ca.G = G_;
    ca.n_surf = n_surf_;
    ca.has_radial = v.has_radial;
    ca.bc_lo[0] = bc_[WEST];
    ca.bc_hi[0] = bc_[EAST];
    ca.bc_lo[1] = bc_[SOUTH];
    ca.bc_hi[1] = bc_[NORTH];
    ca.bc_lo[2] = bc_[BOTTOM];
    ca.bc_hi[2] = bc_[TOP];
    ca.htr = d_htr_;
    ca.cflux = d_cflux_;
    ca.cur = v.cur;
    ca.sfl = v.sfl;
    ca.dtil = d_dtil_;
    ca.stil = d_stil_;
    ca.dhat = d_dhat_;
    ca.shat = d_shat_;
    hipLaunchKernelGGL(k_couple, dim3(nblk((size_t)n_surf_ * G_)), dim3(256), 0, s, ca);
    hipLaunchKernelGGL(k_assemble, dim3(nblk((size_t)nc_ * G_)), dim3(256), 0, s, ca.m, G_, d_hrm_, d_dtil_,
                       d_dhat_, d_coef_);
    SolveArgs a{};
    a.nc = nc_;
    a.nx = nx_;
    a.G = G_;
    a.max_iter = opt_.cmfd_max_iter;
    a.k_tol = opt_.cmfd_k_tol;
    a.psi_tol = opt_.cmfd_psi_tol;
    a.rr = opt_.cmfd_resid_reduction;
    a.coef = d_coef_;
    a.hnf = d_hnf_;
    a.hch = d_hch_;
    a.hsc = d_hsc_;
    a.cvol = d_cvol_;
    a.cflux = d_cflux_;
    a.k = v.k;
    a.info = d_info_;
    const size_t lds = 5 * (size_t)nc_ * sizeof(double);  // buf, fs_old and the BiCGSTAB vectors
    const int kneed = (nc_ + kCmfdThreads - 1) / kCmfdThreads;
    if (kneed <= 1) hipLaunchKernelGGL(k_cmfd_solve<1>, dim3(1), dim3(kCmfdThreads), lds, s, a);
    else if (kneed <= 2) hipLaunchKernelGGL(k_cmfd_solve<2>, dim3(1), dim3(kCmfdThreads), lds, s, a);
    else if (kneed <= 4) hipLaunchKernelGGL(k_cmfd_solve<4>, dim3(1), dim3(kCmfdThreads), lds, s, a);
    else hipLaunchKernelGGL(k_cmfd_solve<6>, dim3(1), dim3(kCmfdThreads), lds, s, a);
    hipLaunchKernelGGL(k_store_currents, dim3(nblk((size_t)n_surf_ * G_)), dim3(256), 0, s, ca, v.cur, v.sfl);
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
    out.assign((size_t)nc_ * G_, 0.0);
    CMFD_CHECK(hipMemcpyAsync(out.data(), d_cflux_, out.size() * 8, hipMemcpyDeviceToHost, s));
    CMFD_CHECK(hipStreamSynchronize(s));
}

}  // namespace gpu
}  // namespace mocc_amd
