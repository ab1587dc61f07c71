// C ABI of the gfx950 MoC sweeper (include/mocgpu.h): device state, the
// source / fission / reduction kernels and the host driver of the sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mocgpu.h"
#include "handle.hpp"

using namespace mocc_amd::gpu;

namespace {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Source::initialize_group + fission + in_scatter (source.cpp:41-112):
// S = chi*fs + sum_{g' != g} Ss[g'->g] phi_g'
__global__ void k_build_source(int F, int G, int g_fixed, int has_fs, const int32_t *fsr_mat,
                               const double *mat_ch, const double *mat_sc, const double *fs,
                               const double *phi_src, double *src)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int f, g;
    if (g_fixed >= 0) {
        f = idx;
        g = g_fixed;
    } else {
        f = idx / G;
        g = idx % G;
    }
    if (f >= F) return;
    const int m = fsr_mat[f];
    double s = 0.0;
    if (has_fs) s += mat_ch[m * G + g] * fs[f];
    const double *row = mat_sc + ((size_t)m * G + g) * G;
    for (int gg = 0; gg < G; gg++) {
        if (gg == g) continue;
        s += row[gg] * phi_src[(size_t)f * G + gg];
    }
    src[(size_t)f * G + g] = s;
}

// TransportSweeper::calc_fission_source (transport_sweeper.cpp:119-134)
__global__ void k_fission_source(int F, int G, const int32_t *fsr_mat, const double *mat_nf, const double *phi,
                                 const double *kptr, double kval, double *fs)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const double rk = 1.0 / (kptr ? kptr[0] : kval);
    const int m = fsr_mat[f];
    double s = 0.0;
    for (int g = 0; g < G; g++) s += rk * mat_nf[m * G + g] * phi[(size_t)f * G + g];
    fs[f] = s;
}



__device__ __forceinline__ void red_values(const RedArgs &a, int mode, int f, double v[4])
{
    v[0] = v[1] = v[2] = v[3] = 0.0;
    const int G = a.G;
    if (mode == RED_FISSION) {
        const int m = a.fsr_mat[f];
        for (int g = 0; g < G; g++) {
            v[0] += a.phi[(size_t)f * G + g] * a.vol[f] * a.mat_nf[m * G + g];
            v[1] += a.phi_old[(size_t)f * G + g] * a.vol[f] * a.mat_nf[m * G + g];
        }
    } else if (mode == RED_NORM) {
        v[0] = a.fs[f] > 0.0 ? 1.0 : 0.0;
        v[1] = a.fs[f];
        v[2] = a.fs_prev[f] > 0.0 ? 1.0 : 0.0;
        v[3] = a.fs_prev[f];
    } else if (mode == RED_PSIERR) {
        const double e = a.fs[f] * a.scal[S_FA] - a.fs_prev[f] * a.scal[S_FB];
        v[0] = e * e;
    } else {
        for (int g = 0; g < G; g++) {
            const double e = a.phi[(size_t)f * G + g] - a.phi_old[(size_t)f * G + g];
            v[0] += e * e;
        }
    }
}

__device__ __forceinline__ void block_sum4(double v[4], double (*sh)[4])
{
    const int t = threadIdx.x;
    for (int q = 0; q < 4; q++) sh[t][q] = v[q];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < 4; q++) sh[t][q] += sh[t + s][q];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRedThreads) void k_red_stage1(RedArgs a, int mode, double *partial)
{
    __shared__ double sh[kRedThreads][4];
    double acc[4] = {0, 0, 0, 0};
    const int stride = gridDim.x * blockDim.x;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < a.F; f += stride) {
        double v[4];
        red_values(a, mode, f, v);
        for (int q = 0; q < 4; q++) acc[q] += v[q];
    }
    block_sum4(acc, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) partial[blockIdx.x * 4 + q] = sh[0][q];
}

// Stage 2 (one block) and the epilogues of EigenSolver::step / solve.
__global__ __launch_bounds__(kRedThreads) void k_red_stage2(RedArgs a, int mode, const double *partial, int nb,
                                                            int update_k, double *out)
{
    __shared__ double sh[kRedThreads][4];
    double acc[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int q = 0; q < 4; q++) acc[q] += partial[b * 4 + q];
    block_sum4(acc, sh);
    if (threadIdx.x != 0) return;
    double *s = a.scal;
    if (mode == RED_FISSION) {
        s[S_TF_NEW] = sh[0][0];
        s[S_TF_OLD] = sh[0][1];
        if (update_k) {
            // k <- k * tfis(new) / tfis(old) (eigen_solver.cpp:242-247)
            s[S_KPREV] = s[S_K];
            s[S_K] = s[S_K] * sh[0][0] / sh[0][1];
            s[S_ERRK] = fabs(s[S_K] - s[S_KPREV]);
        }
    } else if (mode == RED_NORM) {
        // Normalize: scale to sum = number of positive entries (util/utils.hpp:28-51)
        s[S_FA] = sh[0][0] / sh[0][1];
        s[S_FB] = sh[0][2] / sh[0][3];
    } else if (mode == RED_PSIERR) {
        s[S_ERRPSI] = sqrt(sh[0][0] / s[S_NFISS]);
    } else {
        s[S_RESID] = sqrt(sh[0][0]);
    }
    if (out)
        for (int q = 0; q < 4; q++) out[q] = sh[0][q];
}

__global__ void k_copy(const double *a, double *b, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_fill_f(float *a, float v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// MoCSweeper::initialize's state: flat scalar flux 1 (and its old copy), and
// both incoming-flux buffers of every group set to the one-group template.
__global__ void k_init_state(double *phi, double *phi_old, size_t fg, float *psi, const float *tmpl, size_t bcpg,
                             size_t n_psi)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < fg) {
        phi[i] = 1.0;
        phi_old[i] = 1.0;
    }
    if (i < n_psi) psi[i] = tmpl[i % bcpg];
}

// MoCSweeper::get_pin_flux_1g, PIN_PLANE (moc_sweeper.cpp:293-326)
__global__ void k_pin_flux(int npin, int G, int g, const int32_t *pin_off, const int32_t *pin_cell,
                           const double *vol, const double *phi, double *pin)
{
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= npin) return;
    double v = 0.0, pf = 0.0;
    for (int r = pin_off[ip]; r < pin_off[ip + 1]; r++) {
        v += vol[r];
        pf += phi[(size_t)r * G + g] * vol[r];
    }
    pin[pin_cell[ip]] = pf / v;
}

// MoCSweeper::set_pin_flux_1g, PIN_PLANE (moc_sweeper.cpp:396-426): scale the
// pin's fine fluxes by new/old pin average; e^2 per pin for the residual.
__global__ void k_set_pin_flux(int npin, int G, int g, const int32_t *pin_off, const int32_t *pin_cell,
                               const double *area, const double *pin_area, const double *pin_new, double *phi,
                               double *e2)
{
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= npin) return;
    const int ic = pin_cell[ip];
    double fm = 0.0;
    for (int r = pin_off[ip]; r < pin_off[ip + 1]; r++) fm += phi[(size_t)r * G + g] * area[r];
    fm /= pin_area[ip];
    const double e = pin_new[ic] - fm;
    const double fct = pin_new[ic] / fm;
    for (int r = pin_off[ip]; r < pin_off[ip + 1]; r++) phi[(size_t)r * G + g] *= fct;
    e2[ip] = e * e;
}


}  // namespace

// ---------------------------------------------------------------------------
// handle methods that launch kernels
// ---------------------------------------------------------------------------
void mocgpu_handle::reduce(int mode, int update_k)
{
    RedArgs a = red_args();
    hipLaunchKernelGGL(k_red_stage1, dim3(kRedBlocks), dim3(kRedThreads), 0, stream, a, mode, d_partial.p);
    hipLaunchKernelGGL(k_red_stage2, dim3(1), dim3(kRedThreads), 0, stream, a, mode, d_partial.p, kRedBlocks,
                       update_k, d_red_out.p);
}

void mocgpu_handle::build_source(int g_fixed, bool jacobi)
{
    size_t n = g_fixed >= 0 ? (size_t)F : (size_t)F * G;
    hipLaunchKernelGGL(k_build_source, dim3(nblk(n)), dim3(256), 0, stream, F, G, g_fixed, has_fs ? 1 : 0,
                       d_fsr_mat.p, d_mat_ch.p, d_mat_sc.p, d_fs.p, jacobi ? d_phi_old.p : d_phi.p,
                       d_src.p);
}

void mocgpu_handle::launch_sweeps(int g, bool with_current)
{
    SweepArgs a = sweep_args(g);
    if (!opt.gs_boundary) {
        launch_sweep(a, NP, with_current, sweep_grid_items((int)plan->bands.size(), G, g < 0), stream);
        return;
    }
    const int n1 = plan->n_bands_oct1, n2 = (int)plan->bands.size() - n1;
    a.order = d_order_gs.p;
    a.n_bands = n1;
    if (n1 > 0) launch_sweep(a, NP, with_current, sweep_grid_items(n1, G, g < 0), stream);
    a.order = d_order_gs.p + n1;
    a.n_bands = n2;
    if (n2 > 0) launch_sweep(a, NP, with_current, sweep_grid_items(n2, G, g < 0), stream);
}

void mocgpu_handle::sweep_once(int g, bool with_current)
{
    launch_sweeps(g, with_current);
    launch_reduce_flux(reduce_args(g), stream);
    if (with_current) {
        int n0 = nx * ny, n1 = nx * ny + (nx + 1) * ny + (ny + 1) * nx;
        size_t n = (size_t)(n1 - n0) * (g >= 0 ? 1 : G);
        launch_reduce_current(current_args(g), n, stream);
        has_radial = true;
    }
    if (!opt.gs_boundary) {
        if (g >= 0) parity ^= (1ull << g);
        else parity ^= (G >= 64 ? ~0ull : ((1ull << G) - 1ull));
    }
}

void mocgpu_handle::copy_flux_to_old()
{
    size_t n = (size_t)F * G;
    hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(256), 0, stream, d_phi.p, d_phi_old.p, n);
}

namespace {
template <typename F> int guard(mocgpu_handle *h, F &&f)
{
    if (!h) return MOCGPU_ERR_ARG;
    try {
        int rc = f();
        return rc;
    } catch (const HipError &e) {
        if (h) h->err = e.what();
        return MOCGPU_ERR_DEVICE;
    } catch (const std::exception &e) {
        if (h) h->err = e.what();
        return MOCGPU_ERR_ARG;
    }
}

thread_local std::string g_create_err;

// Entry points that launch work on a handle or change its device state refuse
// a handle that a batch owns (from mocgpu_batch_add to mocgpu_batch_remove):
// the batch's launches use the same buffers on the batch's stream.
void require_free(const mocgpu_handle *h, const char *fn)
{
    if (h->batch)
        throw std::runtime_error(std::string(fn) +
                                 ": the handle is a member of a batch (take it out with mocgpu_batch_remove first)");
}

// Reads of a member handle (its flux, boundary fluxes, currents, reductions
// over them) first wait for the work its batch has queued.
void settle(const mocgpu_handle *h)
{
    if (h->batch) batch_sync(h->batch);
}

}  // namespace

extern "C" {

void mocgpu_default_options(mocgpu_options *o)
{
    *o = mocgpu_options{};
    o->n_inner = 1;
    o->energy_mode = 0;
    o->lane_map = 2;
    o->band_rays = 0;
    o->use_cmfd = 0;
    o->cmfd_max_iter = 100;
    o->cmfd_k_tol = 1.0e-6;
    o->cmfd_psi_tol = 1.0e-5;
    o->cmfd_resid_reduction = 0.001;
    o->cmfd_solver = 0;
    o->gs_boundary = 1;
    o->cmfd_negative_fixup = 0;
    o->correction_worker = 0;
}

const char *mocgpu_last_error(const mocgpu_handle *h) { return h ? h->err.c_str() : g_create_err.c_str(); }

}  // extern "C"

namespace {
// mocgpu_create / mocgpu_create_like: with a template handle, the geometry
// (plan, segment streams, exp table, FSR and pin arrays) is shared with it
// and only the cross sections and the state are new.
// The material cross sections of the problem on the handle's stream, and its
// number of fissile FSRs (eigen_solver.cpp:97-111).
void put_materials(mocgpu_handle *h, const mocgpu_problem *p)
{
    const int G = p->n_group;
    const size_t mg = (size_t)p->n_mat * G;
    hipStream_t s = h->stream;
    h->d_mat_tr.refill(std::vector<double>(p->xs_tr, p->xs_tr + mg), s);
    h->d_mat_nf.refill(std::vector<double>(p->xs_nf, p->xs_nf + mg), s);
    h->d_mat_f.refill(std::vector<double>(p->xs_f, p->xs_f + mg), s);
    h->d_mat_ch.refill(std::vector<double>(p->xs_ch, p->xs_ch + mg), s);
    h->d_mat_sc.refill(std::vector<double>(p->xs_sc, p->xs_sc + mg * G), s);
    std::vector<double> ss(mg);
    for (int m = 0; m < p->n_mat; m++)
        for (int g = 0; g < G; g++) ss[(size_t)m * G + g] = p->xs_sc[((size_t)m * G + g) * G + g];
    h->d_mat_ss.refill(ss, s);
    int nfis = 0;
    for (int f = 0; f < p->n_fsr; f++) {
        const int m = h->h_fsr_mat[f];
        for (int g = 0; g < G; g++)
            if (p->xs_nf[(size_t)m * G + g] > 0.0) {
                nfis++;
                break;
            }
    }
    h->n_fissile = nfis;
}

int create_impl(const mocgpu_problem *p, const mocgpu_options *opt, const mocgpu_handle *like,
                mocgpu_handle **out)
{
    if (!p || !out) {
        g_create_err = "mocgpu_create: null problem or output handle";
        return MOCGPU_ERR_ARG;
    }
    auto h = std::make_unique<mocgpu_handle>();
    int rc = guard(h.get(), [&] {
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw HipError("no HIP device");
        if (opt) h->opt = *opt;
        else mocgpu_default_options(&h->opt);
        if (p->n_group > 64) throw std::runtime_error("at most 64 energy groups are supported");
        if (h->opt.lane_map < 0 || h->opt.lane_map > 2)
            throw std::runtime_error("lane_map must be 0, 1 or 2 (-1: autotune)");
        if (like) {
            bool same = p->n_fsr == like->F && p->n_seg == like->n_seg && p->n_polar == like->NP &&
                        p->n_azi == like->A && p->nx == like->nx && p->ny == like->ny &&
                        p->n_ray == (int32_t)like->plan->rays.size();
            for (int i = 0; same && i < 6; i++) same = p->bc[i] == like->bc[i];
            for (int f = 0; same && f < p->n_fsr; f++) same = p->fsr_mat[f] == like->h_fsr_mat[f];
            if (!same) throw std::runtime_error("mocgpu_create_like: the problem's geometry differs from the template's");
            h->plan = like->plan;
            // the shared plan fixes the band width and the lane map
            h->opt.band_rays = like->opt.band_rays;
            h->opt.lane_map = like->opt.lane_map;
        } else {
            h->plan = std::make_shared<SweepPlan>(build_plan(*p, h->opt.band_rays));
        }
        const SweepPlan &pl = *h->plan;
        for (int i = 0; i < 6; i++) h->bc[i] = p->bc[i];
        h->F = p->n_fsr;
        h->G = p->n_group;
        h->NP = p->n_polar;
        h->A = p->n_azi;
        h->nd = pl.nd;
        h->nx = p->nx;
        h->ny = p->ny;
        h->npin = p->nx * p->ny;
        h->n_surf = pl.n_surf;
        h->n_mat = p->n_mat;
        h->n_seg = p->n_seg;
        HIP_CHECK(hipStreamCreate(&h->stream));
        HIP_CHECK(hipEventCreate(&h->ev0));
        HIP_CHECK(hipEventCreate(&h->ev1));
        hipStream_t s = h->stream;
        const int F = h->F, G = h->G;
        if (like) {
            h->d_bands = like->d_bands;
            h->d_order = like->d_order;
            h->d_order_gs = like->d_order_gs;
            h->d_seg_loc = like->d_seg_loc;
            h->d_seg_len = like->d_seg_len;
            h->d_loc_fsr = like->d_loc_fsr;
            h->d_rays = like->d_rays;
            h->d_azi = like->d_azi;
            h->d_ang_rsin = like->d_ang_rsin;
            h->d_ang_wt = like->d_ang_wt;
            h->d_ang_cw = like->d_ang_cw;
            h->d_ang_fw = like->d_ang_fw;
            h->d_slot_ptr = like->d_slot_ptr;
            h->d_slot_pos = like->d_slot_pos;
            h->d_ev_off = like->d_ev_off;
            h->d_ev = like->d_ev;
            h->d_surf_ptr = like->d_surf_ptr;
            h->d_surf_idx = like->d_surf_idx;
            h->d_surf_area = like->d_surf_area;
            h->h_fsr_mat = like->h_fsr_mat;
            h->d_fsr_mat = like->d_fsr_mat;
            h->d_vol = like->d_vol;
            h->d_area = like->d_area;
            h->d_pin_off = like->d_pin_off;
            h->d_pin_cell = like->d_pin_cell;
            h->d_pin_area = like->d_pin_area;
            h->bc_template = like->bc_template;
            h->d_bc_template = like->d_bc_template;
        } else {
        h->d_bands.upload(pl.bands, s);
        h->d_order.upload(pl.band_order, s);
        h->d_order_gs.upload(pl.band_order_gs, s);
        h->d_seg_loc.upload(pl.seg_loc, s);
        h->d_seg_len.upload(pl.seg_len, s);
        h->d_loc_fsr.upload(pl.loc_fsr, s);
        h->d_rays.upload(pl.rays, s);
        h->d_azi.upload(pl.azi, s);
        h->d_ang_rsin.upload(pl.ang_rsin, s);
        h->d_ang_wt.upload(pl.ang_wt, s);
        h->d_ang_cw.upload(pl.ang_cw, s);
        h->d_ang_fw.upload(pl.ang_fw, s);
        h->d_slot_ptr.upload(pl.fsr_slot_ptr, s);
        h->d_slot_pos.upload(pl.slot_pos, s);
        h->d_ev_off.upload(pl.ev_off, s);
        h->d_ev.upload(pl.ev, s);
        h->d_surf_ptr.upload(pl.surf_slot_ptr, s);
        h->d_surf_idx.upload(pl.surf_slot_idx, s);
        h->d_surf_area.upload(pl.surf_area, s);
        h->h_fsr_mat.assign(p->fsr_mat, p->fsr_mat + F);
        h->d_fsr_mat.upload(h->h_fsr_mat, s);
        h->d_vol.upload(std::vector<double>(p->fsr_vol, p->fsr_vol + F), s);
        h->d_area.upload(std::vector<double>(p->fsr_area, p->fsr_area + F), s);
        h->d_pin_off.upload(std::vector<int32_t>(p->pin_fsr_off, p->pin_fsr_off + h->npin + 1), s);
        h->d_pin_cell.upload(std::vector<int32_t>(p->pin_cell, p->pin_cell + h->npin), s);
        h->d_pin_area.upload(std::vector<double>(p->pin_area, p->pin_area + h->npin), s);
        h->sync();
        // the segment streams live on the device from here on
        std::vector<uint16_t>().swap(h->plan->seg_loc);
        std::vector<float>().swap(h->plan->seg_len);
        }
        put_materials(h.get(), p);
        // state
        h->d_phi.alloc((size_t)F * G);
        h->d_phi_old.alloc((size_t)F * G);
        h->d_src.alloc((size_t)F * G);
        h->d_fs.alloc(F);
        h->d_fs_prev.alloc(F);
        h->d_psi.alloc((size_t)G * 2 * pl.bcpg);
        h->d_psum.alloc(pl.loc_fsr.size() * (size_t)G);
        h->d_csum.alloc(std::max<size_t>(1, pl.lsurf_glob.size()) * (size_t)G);
        h->d_ssum.alloc(std::max<size_t>(1, pl.lsurf_glob.size()) * (size_t)G);
        h->d_cur.alloc((size_t)h->n_surf * G);
        h->d_sfl.alloc((size_t)h->n_surf * G);
        h->d_partial.alloc(kRedBlocks * 4);
        h->d_scal.alloc(S_COUNT);
        h->d_red_out.alloc(4);
        h->d_pin.alloc(h->npin);
        h->d_e2.alloc(h->npin);
        HIP_CHECK(hipMemsetAsync(h->d_src.p, 0, sizeof(double) * F * G, s));
        HIP_CHECK(hipMemsetAsync(h->d_fs.p, 0, sizeof(double) * F, s));
        HIP_CHECK(hipMemsetAsync(h->d_cur.p, 0, sizeof(double) * h->n_surf * G, s));
        HIP_CHECK(hipMemsetAsync(h->d_sfl.p, 0, sizeof(double) * h->n_surf * G, s));
        HIP_CHECK(hipMemsetAsync(h->d_psi.p, 0, sizeof(float) * G * 2 * pl.bcpg, s));
        std::vector<double> sc(S_COUNT, 0.0);
        sc[S_K] = 1.0;
        sc[S_KPREV] = 1.0;
        sc[S_NFISS] = (double)h->n_fissile;
        h->d_scal.upload(sc, s);
        // template of the initial incoming flux: 1/(4 pi) on non-vacuum faces
        // (MoCSweeper::initialize + BoundaryCondition::initialize_scalar)
        if (!like) {
        h->bc_template.assign(pl.bcpg, 0.0f);
        const float val = (float)(1.0 / FPI);
        for (int iang = 0; iang < 4 * pl.nd; iang++) {
            int ioct = iang / pl.nd;
            int i1 = iang < 2 * pl.nd ? iang : iang - 2 * pl.nd;
            int a = i1 / pl.NP;
            int ny_ = p->azi_ny[a], nx_ = p->azi_nx[a];
            bool ox_pos = (ioct == 0 || ioct == 3), oy_pos = (ioct == 0 || ioct == 1);
            int sx = ox_pos ? MOCGPU_WEST : MOCGPU_EAST;
            int sy = oy_pos ? MOCGPU_SOUTH : MOCGPU_NORTH;
            if (p->bc[sx] == MOCGPU_BC_REFLECT)
                for (int i = 0; i < ny_; i++) h->bc_template[pl.ang_off[iang] + i] = val;
            if (p->bc[sy] == MOCGPU_BC_REFLECT)
                for (int i = 0; i < nx_; i++) h->bc_template[pl.ang_off[iang] + ny_ + i] = val;
        }
        h->d_bc_template.upload(h->bc_template, s);
        }
        h->d_parity.alloc(1);
        h->d_status.alloc(2);
        h->d_radial.alloc(1);
        h->cmfd = std::make_unique<CmfdSolver>();
        h->cmfd->setup(*p, h->opt, s);
        if (h->opt.use_cmfd || h->opt.correction_worker) {
            h->corr = std::make_unique<CorrSolver>();
            h->corr->setup(*p, pl, s, h->opt.correction_worker != 0);
        }
        h->sync();
        return 0;
    });
    if (rc != 0) {
        g_create_err = h->err;
        return rc;
    }
    *out = h.release();
    return 0;
}
// Autotuner of the ray -> lane mapping: build the plan for each band width,
// time the sweep kernel (HIP events, a few all-group sweeps) with each lane
// map, keep the fastest (band width, lane map) and return its handle reset to
// the initial state.
int autotune_create(const mocgpu_problem *p, const mocgpu_options *opt, mocgpu_handle **out)
{
    mocgpu_options o = *opt;
    const int widths[3] = {64, 48, 32};
    const int maps[3] = {1, 0, 2};
    mocgpu_handle *best = nullptr;
    double best_ms = 0.0;
    int best_map = 1;
    for (int w : widths) {
        o.band_rays = w;
        o.lane_map = 1;
        mocgpu_handle *h = nullptr;
        int rc = create_impl(p, &o, nullptr, &h);
        if (rc == 0) rc = mocgpu_initialize(h);
        if (rc != 0) {
            if (h) mocgpu_destroy(h);
            if (best) mocgpu_destroy(best);
            return rc;
        }
        bool keep = false;
        for (int m : maps) {
            h->opt.lane_map = m;
            double ms = 0.0, kms = 0.0;
            // a first run settles the clocks; the second is the one kept
            rc = mocgpu_time_sweeps(h, 5, &ms, &kms);
            if (rc == 0) rc = mocgpu_time_sweeps(h, 10, &ms, &kms);
            if (rc != 0) {
                mocgpu_destroy(h);
                if (best) mocgpu_destroy(best);
                return rc;
            }
            if (!best || kms < best_ms) {
                best_ms = kms;
                best_map = m;
                keep = true;
                if (best && best != h) mocgpu_destroy(best);
                best = h;
            }
        }
        if (!keep) mocgpu_destroy(h);
    }
    best->opt.lane_map = best_map;
    best->tuned_ms = best_ms;
    int rc = mocgpu_initialize(best);
    if (rc != 0) {
        mocgpu_destroy(best);
        return rc;
    }
    *out = best;
    return 0;
}
}  // namespace

extern "C" {

int mocgpu_create(const mocgpu_problem *p, const mocgpu_options *opt, mocgpu_handle **out)
{
    if (opt && opt->lane_map == -1) return autotune_create(p, opt, out);
    return create_impl(p, opt, nullptr, out);
}

int mocgpu_get_options(const mocgpu_handle *h, mocgpu_options *out)
{
    if (!h || !out) return MOCGPU_ERR_ARG;
    *out = h->opt;
    if (out->band_rays <= 0) out->band_rays = kMaxRaysPerBand;
    return 0;
}

int mocgpu_create_like(const mocgpu_handle *like, const mocgpu_problem *p, const mocgpu_options *opt,
                       mocgpu_handle **out)
{
    if (!like) {
        g_create_err = "mocgpu_create_like: null template handle";
        return MOCGPU_ERR_ARG;
    }
    return create_impl(p, opt, like, out);
}

int64_t mocgpu_state_bytes(const mocgpu_handle *h)
{
    if (!h) return 0;
    const size_t b = h->d_mat_tr.bytes() + h->d_mat_ss.bytes() + h->d_mat_nf.bytes() + h->d_mat_ch.bytes() +
                     h->d_mat_sc.bytes() + h->d_mat_f.bytes() + h->d_phi.bytes() + h->d_phi_old.bytes() +
                     h->d_src.bytes() + h->d_fs.bytes() + h->d_fs_prev.bytes() + h->d_psum.bytes() +
                     h->d_csum.bytes() + h->d_ssum.bytes() + h->d_cur.bytes() + h->d_sfl.bytes() +
                     h->d_psi.bytes() + h->d_partial.bytes() + h->d_scal.bytes() + h->d_red_out.bytes() +
                     h->d_pin.bytes() + h->d_e2.bytes() + h->d_parity.bytes() + h->d_status.bytes() +
                     h->d_radial.bytes();
    return (int64_t)b;
}

int64_t mocgpu_geometry_bytes(const mocgpu_handle *h)
{
    if (!h) return 0;
    const size_t b = h->d_bands.bytes() + h->d_order.bytes() + h->d_order_gs.bytes() + h->d_seg_loc.bytes() + h->d_seg_len.bytes() +
                     h->d_loc_fsr.bytes() + h->d_rays.bytes() + h->d_azi.bytes() + h->d_ang_rsin.bytes() +
                     h->d_ang_wt.bytes() + h->d_ang_cw.bytes() + h->d_ang_fw.bytes() + h->d_slot_ptr.bytes() +
                     h->d_slot_pos.bytes() + h->d_ev_off.bytes() + h->d_ev.bytes() + h->d_surf_ptr.bytes() +
                     h->d_surf_idx.bytes() + h->d_surf_area.bytes() + h->d_fsr_mat.bytes() +
                     h->d_vol.bytes() + h->d_area.bytes() + h->d_pin_off.bytes() + h->d_pin_cell.bytes() +
                     h->d_pin_area.bytes() + (h->corr ? h->corr->bytes() : 0);
    return (int64_t)b;
}

int mocgpu_destroy(mocgpu_handle *h)
{
    // a member handle leaves its batch first, so the batch never launches
    // on freed buffers
    if (h && h->batch) batch_detach(h->batch, h);
    delete h;
    return 0;
}

int mocgpu_reload_xs(mocgpu_handle *h, const mocgpu_problem *p)
{
    if (!p) return MOCGPU_ERR_ARG;
    return guard(h, [&] {
        require_free(h, "mocgpu_reload_xs");
        bool same = p->n_fsr == h->F && p->n_group == h->G && p->n_mat == h->n_mat && p->n_seg == h->n_seg &&
                    p->n_polar == h->NP && p->n_azi == h->A && p->nx == h->nx && p->ny == h->ny;
        for (int f = 0; same && f < p->n_fsr; f++) same = p->fsr_mat[f] == h->h_fsr_mat[f];
        if (!same)
            throw std::runtime_error("mocgpu_reload_xs: the problem's geometry or group count differs from the handle's");
        put_materials(h, p);
        if (h->cmfd) h->cmfd->update_xs(*p, h->stream);
        // the state a new handle starts from (mocgpu_initialize sets the rest)
        const size_t fg = (size_t)h->F * h->G;
        HIP_CHECK(hipMemsetAsync(h->d_src.p, 0, sizeof(double) * fg, h->stream));
        HIP_CHECK(hipMemsetAsync(h->d_fs.p, 0, sizeof(double) * h->F, h->stream));
        std::vector<double> sc(S_COUNT, 0.0);
        sc[S_K] = 1.0;
        sc[S_KPREV] = 1.0;
        sc[S_NFISS] = (double)h->n_fissile;
        HIP_CHECK(hipMemcpyAsync(h->d_scal.p, sc.data(), sc.size() * 8, hipMemcpyHostToDevice, h->stream));
        h->sync();
        return mocgpu_initialize(h);
    });
}

int mocgpu_initialize(mocgpu_handle *h)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_initialize");
        const size_t fg = (size_t)h->F * h->G, bcpg = (size_t)h->plan->bcpg, n_psi = (size_t)h->G * 2 * bcpg;
        const size_t n = std::max(fg, n_psi);
        hipLaunchKernelGGL(k_init_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->d_phi.p,
                           h->d_phi_old.p, fg, h->d_psi.p, h->d_bc_template.p, bcpg, n_psi);
        HIP_CHECK(hipGetLastError());
        h->parity = 0;
        // no coarse-surface currents until a sweep tallies them: the first
        // CMFD solve after initialize runs with D-hat = 0, as after
        // mocgpu_batch_add
        h->has_radial = false;
        if (h->d_cur.p) HIP_CHECK(hipMemsetAsync(h->d_cur.p, 0, h->d_cur.bytes(), h->stream));
        if (h->d_sfl.p) HIP_CHECK(hipMemsetAsync(h->d_sfl.p, 0, h->d_sfl.bytes(), h->stream));
        const int32_t no_radial = 0;
        if (h->d_radial.p) HIP_CHECK(hipMemcpyAsync(h->d_radial.p, &no_radial, 4, hipMemcpyHostToDevice, h->stream));
        h->cmfd->initialize(h->stream);
        h->sync();
        return 0;
    });
}

int mocgpu_set_flux(mocgpu_handle *h, const double *phi)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_set_flux");
        HIP_CHECK(hipMemcpyAsync(h->d_phi.p, phi, sizeof(double) * h->F * h->G, hipMemcpyHostToDevice, h->stream));
        h->sync();
        return 0;
    });
}

int mocgpu_get_flux(mocgpu_handle *h, double *phi)
{
    return guard(h, [&] {
        settle(h);
        HIP_CHECK(hipMemcpyAsync(phi, h->d_phi.p, sizeof(double) * h->F * h->G, hipMemcpyDeviceToHost, h->stream));
        h->sync();
        return 0;
    });
}

int64_t mocgpu_boundary_size(const mocgpu_handle *h) { return h ? (int64_t)h->G * h->plan->bcpg : 0; }

int mocgpu_get_boundary(mocgpu_handle *h, double *bc, int64_t n)
{
    return guard(h, [&] {
        settle(h);
        if (n != (int64_t)h->G * h->plan->bcpg) throw std::runtime_error("boundary size mismatch");
        std::vector<float> tmp(h->plan->bcpg);
        for (int g = 0; g < h->G; g++) {
            int par = (int)((h->parity >> g) & 1ull);
            HIP_CHECK(hipMemcpyAsync(tmp.data(), h->d_psi.p + ((size_t)g * 2 + par) * h->plan->bcpg, tmp.size() * 4,
                                     hipMemcpyDeviceToHost, h->stream));
            h->sync();
            for (int i = 0; i < h->plan->bcpg; i++) bc[(size_t)g * h->plan->bcpg + i] = tmp[i];
        }
        return 0;
    });
}

int mocgpu_set_boundary(mocgpu_handle *h, const double *bc, int64_t n)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_set_boundary");
        if (n != (int64_t)h->G * h->plan->bcpg) throw std::runtime_error("boundary size mismatch");
        std::vector<float> tmp(h->plan->bcpg);
        for (int g = 0; g < h->G; g++) {
            for (int i = 0; i < h->plan->bcpg; i++) tmp[i] = (float)bc[(size_t)g * h->plan->bcpg + i];
            int par = (int)((h->parity >> g) & 1ull);
            HIP_CHECK(hipMemcpyAsync(h->d_psi.p + ((size_t)g * 2 + par) * h->plan->bcpg, tmp.data(), tmp.size() * 4,
                                     hipMemcpyHostToDevice, h->stream));
            h->sync();
        }
        return 0;
    });
}

int mocgpu_store_old_flux(mocgpu_handle *h)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_store_old_flux");
        h->copy_flux_to_old();
        h->sync();
        return 0;
    });
}

int mocgpu_calc_fission_source(mocgpu_handle *h, double k, double *fs_out)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_calc_fission_source");
        hipLaunchKernelGGL(k_fission_source, dim3(nblk(h->F)), dim3(256), 0, h->stream, h->F, h->G,
                           h->d_fsr_mat.p, h->d_mat_nf.p, h->d_phi.p, (const double *)nullptr, k, h->d_fs.p);
        h->has_fs = true;
        if (fs_out)
            HIP_CHECK(hipMemcpyAsync(fs_out, h->d_fs.p, sizeof(double) * h->F, hipMemcpyDeviceToHost, h->stream));
        h->sync();
        return 0;
    });
}

int mocgpu_set_fission_source(mocgpu_handle *h, const double *fs)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_set_fission_source");
        HIP_CHECK(hipMemcpyAsync(h->d_fs.p, fs, sizeof(double) * h->F, hipMemcpyHostToDevice, h->stream));
        h->has_fs = true;
        h->sync();
        return 0;
    });
}

int mocgpu_build_source(mocgpu_handle *h, int32_t g)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_build_source");
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        h->build_source(g, h->opt.energy_mode == 1);
        h->sync();
        return 0;
    });
}

int mocgpu_sweep(mocgpu_handle *h, int32_t g, int32_t with_current)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_sweep");
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        h->mocsweep(g, with_current != 0);
        h->sync();
        return 0;
    });
}

int mocgpu_sweep_all(mocgpu_handle *h, int32_t with_current)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_sweep_all");
        h->sweep_once(-1, with_current != 0);
        return 0;
    });
}

int mocgpu_synchronize(mocgpu_handle *h)
{
    return guard(h, [&] {
        h->sync();
        return 0;
    });
}

int mocgpu_fss_step(mocgpu_handle *h, int32_t with_current)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_fss_step");
        h->fss_step(with_current != 0);
        h->sync();
        return 0;
    });
}

int mocgpu_total_fission(mocgpu_handle *h, int32_t old, double *out)
{
    return guard(h, [&] {
        settle(h);
        h->reduce(RED_FISSION, 0);
        double v[4];
        HIP_CHECK(hipMemcpyAsync(v, h->d_red_out.p, sizeof(v), hipMemcpyDeviceToHost, h->stream));
        h->sync();
        *out = old ? v[1] : v[0];
        return 0;
    });
}

int mocgpu_flux_residual(mocgpu_handle *h, double *out)
{
    return guard(h, [&] {
        settle(h);
        h->reduce(RED_RESID, 0);
        double v[4];
        HIP_CHECK(hipMemcpyAsync(v, h->d_red_out.p, sizeof(v), hipMemcpyDeviceToHost, h->stream));
        h->sync();
        *out = std::sqrt(v[0]);
        return 0;
    });
}

int mocgpu_get_pin_flux_1g(mocgpu_handle *h, int32_t g, double *pin)
{
    return guard(h, [&] {
        settle(h);
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        if (!pin) throw std::runtime_error("null pin flux array");
        hipLaunchKernelGGL(k_pin_flux, dim3(nblk(h->npin)), dim3(256), 0, h->stream, h->npin, h->G, g,
                           h->d_pin_off.p, h->d_pin_cell.p, h->d_vol.p, h->d_phi.p, h->d_pin.p);
        HIP_CHECK(hipMemcpyAsync(pin, h->d_pin.p, sizeof(double) * h->npin, hipMemcpyDeviceToHost, h->stream));
        h->sync();
        return 0;
    });
}

int mocgpu_set_pin_flux_1g(mocgpu_handle *h, int32_t g, const double *pin, double *resid)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_set_pin_flux_1g");
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        if (!pin) throw std::runtime_error("null pin flux array");
        // moc_sweeper.cpp:367-375 throws on a non-positive input flux
        for (int i = 0; i < h->npin; i++)
            if (!(pin[i] > 0.0)) {
                h->err = "Negative or zero input flux: " + std::to_string(pin[i]);
                return (int)MOCGPU_ERR_NONPOS_FLUX;
            }
        HIP_CHECK(hipMemcpyAsync(h->d_pin.p, pin, sizeof(double) * h->npin, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_set_pin_flux, dim3(nblk(h->npin)), dim3(256), 0, h->stream, h->npin, h->G, g,
                           h->d_pin_off.p, h->d_pin_cell.p, h->d_area.p, h->d_pin_area.p, h->d_pin.p, h->d_phi.p,
                           h->d_e2.p);
        std::vector<double> e2(h->npin);
        HIP_CHECK(hipMemcpyAsync(e2.data(), h->d_e2.p, sizeof(double) * h->npin, hipMemcpyDeviceToHost, h->stream));
        h->sync();
        double r = 0.0;
        for (double v : e2) r += v;
        if (resid) *resid = std::sqrt(r) / h->npin;
        return 0;
    });
}

int32_t mocgpu_n_surf(const mocgpu_handle *h) { return h ? h->n_surf : 0; }

int mocgpu_get_currents(mocgpu_handle *h, int32_t g, double *current, double *surface_flux)
{
    return guard(h, [&] {
        settle(h);
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        std::vector<double> c((size_t)h->n_surf * h->G), s((size_t)h->n_surf * h->G);
        HIP_CHECK(hipMemcpyAsync(c.data(), h->d_cur.p, c.size() * 8, hipMemcpyDeviceToHost, h->stream));
        HIP_CHECK(hipMemcpyAsync(s.data(), h->d_sfl.p, s.size() * 8, hipMemcpyDeviceToHost, h->stream));
        h->sync();
        for (int i = 0; i < h->n_surf; i++) {
            if (current) current[i] = c[(size_t)i * h->G + g];
            if (surface_flux) surface_flux[i] = s[(size_t)i * h->G + g];
        }
        return 0;
    });
}

int mocgpu_set_currents(mocgpu_handle *h, int32_t g, const double *current, const double *surface_flux)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_set_currents");
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        std::vector<double> c((size_t)h->n_surf * h->G), s((size_t)h->n_surf * h->G);
        HIP_CHECK(hipMemcpyAsync(c.data(), h->d_cur.p, c.size() * 8, hipMemcpyDeviceToHost, h->stream));
        HIP_CHECK(hipMemcpyAsync(s.data(), h->d_sfl.p, s.size() * 8, hipMemcpyDeviceToHost, h->stream));
        h->sync();
        for (int i = 0; i < h->n_surf; i++) {
            if (current) c[(size_t)i * h->G + g] = current[i];
            if (surface_flux) s[(size_t)i * h->G + g] = surface_flux[i];
        }
        HIP_CHECK(hipMemcpyAsync(h->d_cur.p, c.data(), c.size() * 8, hipMemcpyHostToDevice, h->stream));
        HIP_CHECK(hipMemcpyAsync(h->d_sfl.p, s.data(), s.size() * 8, hipMemcpyHostToDevice, h->stream));
        h->sync();
        h->has_radial = true;
        return 0;
    });
}

int32_t mocgpu_attenuation(const float *tau, float *out, int32_t n)
{
    if (n <= 0) return 0;
    if (!tau || !out) return MOCGPU_ERR_ARG;
    float *d = nullptr;
    if (hipMalloc(&d, 2 * sizeof(float) * (size_t)n) != hipSuccess) return MOCGPU_ERR_DEVICE;
    hipError_t e = hipMemcpy(d, tau, sizeof(float) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        mocc_amd::gpu::launch_attenuation(d, d + n, n, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d + n, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? 0 : MOCGPU_ERR_DEVICE;
}

int64_t mocgpu_updates_per_sweep(const mocgpu_handle *h) { return h ? 2LL * h->NP * h->n_seg * h->G : 0; }

int mocgpu_time_sweeps(mocgpu_handle *h, int32_t n, double *ms_per_sweep, double *kernel_ms)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_time_sweeps");
        if (n <= 0) throw std::runtime_error("n_sweeps must be positive");
        // the launches of sweep_once(-1, false), with events around the sweep kernels
        auto launch = [&](bool timed) {
            if (timed) HIP_CHECK(hipEventRecord(h->ev0, h->stream));
            h->launch_sweeps(-1, false);
            if (!h->opt.gs_boundary) h->parity ^= (h->G >= 64 ? ~0ull : ((1ull << h->G) - 1ull));
            if (timed) HIP_CHECK(hipEventRecord(h->ev1, h->stream));
            launch_reduce_flux(h->reduce_args(-1), h->stream);
        };
        launch(false);  // warm-up
        h->sync();
        float ksum = 0.0f;
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < n; i++) {
            launch(true);
            HIP_CHECK(hipEventSynchronize(h->ev1));
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
            ksum += ms;
        }
        h->sync();
        auto t1 = std::chrono::steady_clock::now();
        *ms_per_sweep = std::chrono::duration<double, std::milli>(t1 - t0).count() / n;
        *kernel_ms = ksum / n;
        return 0;
    });
}

// EigenSolver::step + convergence bookkeeping (eigen_solver.cpp:162-253),
// device-resident: k lives in device memory, the host reads the scalars back.
int mocgpu_eigen_iterate(mocgpu_handle *h, mocgpu_eigen_state *st)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_eigen_iterate");
        HIP_CHECK(hipMemcpyAsync(h->d_scal.p + S_K, &st->k, 8, hipMemcpyHostToDevice, h->stream));
        if (h->opt.use_cmfd) h->cmfd->enqueue_step(h->cmfd_view(), h->stream);
        hipLaunchKernelGGL(k_fission_source, dim3(nblk(h->F)), dim3(256), 0, h->stream, h->F, h->G, h->d_fsr_mat.p,
                           h->d_mat_nf.p, h->d_phi.p, h->d_scal.p + S_K, 0.0, h->d_fs.p);
        h->has_fs = true;
        hipLaunchKernelGGL(k_copy, dim3(nblk(h->F)), dim3(256), 0, h->stream, h->d_fs.p, h->d_fs_prev.p,
                           (size_t)h->F);
        h->fss_step(h->opt.use_cmfd != 0);
        h->reduce(RED_FISSION, 1);
        hipLaunchKernelGGL(k_fission_source, dim3(nblk(h->F)), dim3(256), 0, h->stream, h->F, h->G, h->d_fsr_mat.p,
                           h->d_mat_nf.p, h->d_phi.p, h->d_scal.p + S_K, 0.0, h->d_fs.p);
        h->reduce(RED_NORM, 0);
        h->reduce(RED_PSIERR, 0);
        double sc[S_COUNT];
        HIP_CHECK(hipMemcpyAsync(sc, h->d_scal.p, sizeof(sc), hipMemcpyDeviceToHost, h->stream));
        h->sync();
        if (h->opt.use_cmfd) {
            int bad = 0;
            h->cmfd->read_info(h->stream, &h->cmfd_iterations, &bad, &h->cmfd_bicgstab);
            if (bad) {
                h->err = "Negative or zero input flux in the CMFD rescale";
                return MOCGPU_ERR_NONPOS_FLUX;
            }
        }
        st->k = sc[S_K];
        st->k_prev = sc[S_KPREV];
        st->err_k = sc[S_ERRK];
        st->err_psi = sc[S_ERRPSI];
        st->iterations++;
        if (st->k != st->k) {
            h->err = "Eigenvalue is not a number. Giving up.";
            return MOCGPU_ERR_NAN;
        }
        return 0;
    });
}

// EigenSolver::solve (eigen_solver.cpp:131-223)
int mocgpu_eigen_solve(mocgpu_handle *h, double k_tol, double psi_tol, int32_t max_iter, int32_t min_iter,
                       mocgpu_eigen_state *st)
{
    return mocgpu_eigen_solve_history(h, k_tol, psi_tol, max_iter, min_iter, st, nullptr, 0);
}

// The start of EigenSolver::solve (eigen_solver.cpp:131-160): the sweeper's
// initial state, k = 1, the iteration counters cleared.
int mocgpu_eigen_begin(mocgpu_handle *h, mocgpu_eigen_state *st)
{
    if (!st) return MOCGPU_ERR_ARG;
    int rc = mocgpu_initialize(h);
    if (rc) return rc;
    *st = mocgpu_eigen_state{};
    st->k = 1.0;
    st->k_prev = 1.0;
    return guard(h, [&] {
        std::vector<double> sc(S_COUNT, 0.0);
        sc[S_K] = 1.0;
        sc[S_KPREV] = 1.0;
        sc[S_NFISS] = (double)h->n_fissile;
        HIP_CHECK(hipMemcpyAsync(h->d_scal.p, sc.data(), sc.size() * 8, hipMemcpyHostToDevice, h->stream));
        h->sync();
        return 0;
    });
}

int mocgpu_eigen_solve_history(mocgpu_handle *h, double k_tol, double psi_tol, int32_t max_iter,
                               int32_t min_iter, mocgpu_eigen_state *st, double *history, int32_t cap)
{
    if (!st || (cap > 0 && !history)) return MOCGPU_ERR_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = mocgpu_eigen_begin(h, st);
    if (rc) return rc;
    for (int n = 0; n < max_iter; n++) {
        rc = mocgpu_eigen_iterate(h, st);
        if (rc) return rc;
        if (n < cap) {
            double *row = history + 4 * (size_t)n;
            row[0] = st->k;
            row[1] = st->err_k;
            row[2] = st->err_psi;
            row[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        if (st->err_k < k_tol && st->err_psi < psi_tol && n >= min_iter) {
            st->converged = 1;
            return 0;
        }
    }
    h->err = "Maximum number of iterations reached!";
    return MOCGPU_ERR_MAXITER;
}

int mocgpu_cmfd_step(mocgpu_handle *h, double *k)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_cmfd_step");
        if (!h->opt.use_cmfd) throw std::runtime_error("CMFD is not enabled (options.use_cmfd = 0)");
        HIP_CHECK(hipMemcpyAsync(h->d_scal.p + S_K, k, 8, hipMemcpyHostToDevice, h->stream));
        int rc = h->cmfd_step_sync(&h->cmfd_iterations);
        if (rc) return rc;
        HIP_CHECK(hipMemcpyAsync(k, h->d_scal.p + S_K, 8, hipMemcpyDeviceToHost, h->stream));
        h->sync();
        return 0;
    });
}

int32_t mocgpu_cmfd_iterations(const mocgpu_handle *h) { return h ? h->cmfd_iterations : 0; }

int32_t mocgpu_cmfd_bicgstab_iterations(const mocgpu_handle *h) { return h ? h->cmfd_bicgstab : 0; }

int mocgpu_cmfd_flux(mocgpu_handle *h, double *out)
{
    return guard(h, [&] {
        settle(h);
        if (!h->opt.use_cmfd) throw std::runtime_error("CMFD is not enabled (options.use_cmfd = 0)");
        std::vector<double> f;
        h->cmfd->coarse_flux(h->stream, f);
        std::copy(f.begin(), f.end(), out);
        return 0;
    });
}

int mocgpu_cmfd_matrix(mocgpu_handle *h, int32_t g, int32_t *row_ptr, int32_t *col, double *val, int32_t cap)
{
    return guard(h, [&] {
        settle(h);
        if (!h->opt.use_cmfd) throw std::runtime_error("CMFD is not enabled (options.use_cmfd = 0)");
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        std::vector<int32_t> rp, cl;
        std::vector<double> vl;
        h->cmfd->matrix_csr(g, h->stream, rp, cl, vl);
        if ((int)cl.size() > cap) return (int)MOCGPU_ERR_CAPACITY;
        std::copy(rp.begin(), rp.end(), row_ptr);
        std::copy(cl.begin(), cl.end(), col);
        std::copy(vl.begin(), vl.end(), val);
        return (int)cl.size();
    });
}

// MoCSweeper_2D3D::sweep (moc_sweeper_2d3d.cpp:48-99): the inner iterations
// of group g, the last one with the CurrentCorrections worker. The pin-cell
// cross sections are homogenized from the flux before that sweep; its band
// sweep updates the flux and boundary as any sweep, then the correction sweep
// (from a copy of the same inputs) tallies the factors and the coarse currents.
int mocgpu_sweep_2d3d(mocgpu_handle *h, int32_t g)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_sweep_2d3d");
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        if (!h->corr || !h->corr->full())
            throw std::runtime_error("the handle was created without the correction worker "
                                     "(options.correction_worker = 0)");
        if (!h->opt.use_cmfd)
            throw std::runtime_error("2D3D MoC sweeper needs coarse data to collect calculate correction "
                                     "factors. Try enabling CMFD.");
        for (int inner = 0; inner < h->opt.n_inner; inner++) {
            if (inner == h->opt.n_inner - 1) {
                CmfdView v = h->cmfd_view();
                h->cmfd->homogenize(v, h->stream);
                CorrView cv{};
                cv.F = h->F;
                cv.G = h->G;
                cv.g = g;
                cv.bcpg = h->plan->bcpg;
                cv.gs = h->opt.gs_boundary ? 1 : 0;
                cv.parity = (int)((h->parity >> g) & 1ull) << g;
                cv.psi = h->d_psi.p;
                cv.src = h->d_src.p;
                cv.phi = h->d_phi.p;
                cv.fsr_mat = h->d_fsr_mat.p;
                cv.mat_tr = h->d_mat_tr.p;
                cv.mat_ss = h->d_mat_ss.p;
                cv.htr = h->cmfd->homog_tr();
                cv.cur = h->d_cur.p;
                cv.sfl = h->d_sfl.p;
                h->corr->before_sweep(cv, h->stream);
                h->sweep_once(g, false);
                h->corr->after_sweep(cv, h->stream);
                h->has_radial = true;
            } else {
                h->sweep_once(g, false);
            }
        }
        h->sync();
        return 0;
    });
}

int mocgpu_get_corrections(mocgpu_handle *h, double *alpha, double *beta)
{
    return guard(h, [&] {
        settle(h);
        if (!h->corr || !h->corr->full())
            throw std::runtime_error("the handle was created without the correction worker");
        h->corr->corrections(h->stream, alpha, beta);
        return 0;
    });
}

int mocgpu_correction_residual(mocgpu_handle *h, int32_t g, double *out)
{
    return guard(h, [&] {
        settle(h);
        if (g < 0 || g >= h->G) throw std::runtime_error("group out of range");
        if (!h->corr || !h->corr->full())
            throw std::runtime_error("the handle was created without the correction worker");
        h->corr->residual(h->stream, g, out);
        return 0;
    });
}

// MoCSweeper::update_incoming_flux (moc_sweeper.cpp:255-291) from the partial
// currents the last CMFD solves stored.
int mocgpu_update_incoming_flux(mocgpu_handle *h)
{
    return guard(h, [&] {
        require_free(h, "mocgpu_update_incoming_flux");
        if (!h->opt.use_cmfd || !h->corr)
            throw std::runtime_error("update_incoming_flux needs the CMFD partial currents (options.use_cmfd = 0)");
        h->corr->update_incoming(h->d_psi.p, h->parity, h->G, h->cmfd->partial(), h->cmfd->partial_old(),
                                 h->cmfd->has_old_partial(), h->stream);
        h->sync();
        return 0;
    });
}

}  // extern "C"
