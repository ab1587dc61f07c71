// Device state of one problem: the mocgpu_handle behind the C ABI, shared by
// the single-problem entry points (mocgpu.hip) and the batch engine
// (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mocgpu.h"
#include "cmfd.hpp"
#include "corr.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace mocc_amd {
namespace gpu {

constexpr double PI = 3.1415926535897932;
constexpr double FPI = 4.0 * PI;
constexpr int kRedBlocks = 256;  // fixed stage-1 grid: reductions are bitwise reproducible
constexpr int kRedThreads = 256;

struct HipError : std::runtime_error {
    explicit HipError(const std::string &m) : std::runtime_error(m) {}
};

#define HIP_CHECK(expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

// Device buffer with shared ownership: copying a DBuf shares the allocation
// (handles of one geometry share their plan arrays); the memory is freed when
// the last copy goes.
template <typename T> struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    std::shared_ptr<void> mem;
    void alloc(size_t count)
    {
        free();
        n = count;
        if (count) {
            void *q = nullptr;
            HIP_CHECK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
            mem = std::shared_ptr<void>(q, [](void *x) { (void)hipFree(x); });
            p = (T *)q;
        }
    }
    void upload(const std::vector<T> &v, hipStream_t s)
    {
        alloc(v.size());
        if (!v.empty()) HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    // upload into the present allocation when it has the size and this
    // buffer alone owns it (no allocation, no free); otherwise as upload
    void refill(const std::vector<T> &v, hipStream_t s)
    {
        if (!p || n != v.size() || mem.use_count() != 1) return upload(v, s);
        if (!v.empty()) HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void free()
    {
        mem.reset();
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

enum RedMode { RED_FISSION = 0, RED_NORM = 1, RED_PSIERR = 2, RED_RESID = 3 };

struct RedArgs {
    int F, G;
    const int32_t *fsr_mat;
    const double *mat_nf, *vol, *phi, *phi_old, *fs, *fs_prev;
    double *scal;  // device scalars (see Scal)
};

// device scalar slots
enum Scal { S_K = 0, S_KPREV, S_TF_NEW, S_TF_OLD, S_ERRK, S_ERRPSI, S_FA, S_FB, S_NFISS, S_RESID, S_RESERVED, S_COUNT };

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace gpu
}  // namespace mocc_amd

using namespace mocc_amd::gpu;  // the handle is the C ABI's global type; its members use the library's types

// Take h out of batch b (mocgpu_destroy of a member handle), and wait for the
// work batch b has queued; defined with the batch engine.
void batch_detach(mocgpu_batch *b, struct mocgpu_handle *h);
void batch_sync(mocgpu_batch *b);

struct mocgpu_handle {
    std::string err;
    mocgpu_options opt{};
    // The batch this handle is a member of (mocgpu_batch_add), or null. While
    // it is set, the batch's launches own the handle's device state.
    mocgpu_batch *batch = nullptr;
    std::shared_ptr<SweepPlan> plan;  // host plan, shared by the handles of one geometry
    int F = 0, G = 0, NP = 0, A = 0, nd = 0, nx = 0, ny = 0, npin = 0, n_surf = 0, n_mat = 0;
    int n_fissile = 0;
    int bc[6] = {0, 0, 0, 0, 0, 0};
    int64_t n_seg = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t parity = 0;
    bool has_fs = false;
    bool has_radial = false;
    int cmfd_iterations = 0, cmfd_bicgstab = 0;
    double tuned_ms = 0.0;  // sweep kernel time of the autotuned lane map

    // plan
    DBuf<BandInfo> d_bands;
    DBuf<int32_t> d_order, d_order_gs, d_loc_fsr, d_slot_ptr, d_slot_pos, d_ev_off, d_surf_ptr, d_surf_idx;
    DBuf<uint16_t> d_seg_loc;
    DBuf<float> d_seg_len, d_ang_rsin, d_ang_wt, d_ang_cw, d_ang_fw;
    DBuf<RayInfo> d_rays;
    DBuf<AziInfo> d_azi;
    DBuf<uint32_t> d_ev;
    DBuf<double> d_surf_area;
    // cross sections and mesh
    DBuf<int32_t> d_fsr_mat, d_pin_off, d_pin_cell;
    DBuf<double> d_mat_tr, d_mat_ss, d_mat_nf, d_mat_ch, d_mat_sc, d_mat_f, d_vol, d_area, d_pin_area;
    // state
    DBuf<double> d_phi, d_phi_old, d_src, d_fs, d_fs_prev, d_psum, d_csum, d_ssum, d_cur, d_sfl;
    DBuf<float> d_psi;
    DBuf<double> d_partial, d_scal, d_red_out, d_pin, d_e2;
    // batch mode: device parity word, (iterations, finished) status and the
    // flag that radial currents exist (read by the CMFD coupling in replays)
    DBuf<uint64_t> d_parity;
    DBuf<int32_t> d_status, d_radial;
    std::unique_ptr<CmfdSolver> cmfd;
    // rays for update_incoming_flux (CMFD handles) and the 2-D/3-D
    // correction sweep (options.correction_worker)
    std::unique_ptr<CorrSolver> corr;

    // host copies
    std::vector<float> bc_template;  // one group of incoming values after initialize
    DBuf<float> d_bc_template;        // the same on the device (shared with create_like handles)
    std::vector<int32_t> h_fsr_mat;

    ~mocgpu_handle()
    {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    SweepArgs sweep_args(int g_fixed) const
    {
        SweepArgs a{};
        a.bands = d_bands.p;
        a.order = d_order.p;
        a.n_bands = (int)plan->bands.size();
        a.g_fixed = g_fixed;
        a.G = G;
        a.F = F;
        a.bcpg = plan->bcpg;
        a.lds_local = plan->max_local + kDummySlots;
        a.lds_lsurf = std::max(1, plan->max_lsurf);
        a.n_slots = (int)plan->loc_fsr.size();
        a.n_lslots = (int)std::max<size_t>(1, plan->lsurf_glob.size());
        a.parity = parity;
        a.gs = opt.gs_boundary ? 1 : 0;
        a.interleave = opt.lane_map;  // 0, 1 or 2 (checked at create)
        a.wg_rays = plan->max_band_rays <= 32 ? 32 : kRays;
        a.seg_loc = d_seg_loc.p;
        a.seg_len = d_seg_len.p;
        a.loc_fsr = d_loc_fsr.p;
        a.rays = d_rays.p;
        a.azi = d_azi.p;
        a.ang_rsin = d_ang_rsin.p;
        a.ang_wt = d_ang_wt.p;
        a.ang_cw = d_ang_cw.p;
        a.ang_fw = d_ang_fw.p;
        a.fsr_mat = d_fsr_mat.p;
        a.mat_tr = d_mat_tr.p;
        a.mat_ss = d_mat_ss.p;
        a.src = d_src.p;
        a.phi = d_phi.p;
        a.psi = d_psi.p;
        a.psum = d_psum.p;
        psum_strides(g_fixed, a.psum_sp, a.psum_sg);
        a.slot_pos = d_slot_pos.p;
        a.ev_off = d_ev_off.p;
        a.ev = d_ev.p;
        a.csum = d_csum.p;
        a.ssum = d_ssum.p;
        return a;
    }

    // Layout of the band partials for a sweep of group g (g < 0: all groups)
    // and its reduction. All groups: position-major, so the G workgroups of a
    // band (one XCD) fill whole lines of a position's groups. One group:
    // group-major, so its reduction reads only that group's partials.
    void psum_strides(int g, int64_t &sp, int64_t &sg) const
    {
        if (g < 0) {
            sp = G;
            sg = 1;
        } else {
            sp = 1;
            sg = (int64_t)plan->loc_fsr.size();
        }
    }

    ReduceArgs reduce_args(int g_fixed) const
    {
        ReduceArgs r{};
        r.F = F;
        r.G = G;
        r.g_fixed = g_fixed;
        r.n_slots = (int)plan->loc_fsr.size();
        r.slot_ptr = d_slot_ptr.p;
        r.psum = d_psum.p;
        psum_strides(g_fixed, r.psum_sp, r.psum_sg);
        r.fsr_mat = d_fsr_mat.p;
        r.mat_tr = d_mat_tr.p;
        r.mat_ss = d_mat_ss.p;
        r.vol = d_vol.p;
        r.src = d_src.p;
        r.phi = d_phi.p;
        return r;
    }

    // coarse-surface current reduction of the radial surfaces (g < 0: all groups)
    CurrentArgs current_args(int g_fixed) const
    {
        CurrentArgs c{};
        c.n0 = nx * ny;
        c.n1 = nx * ny + (nx + 1) * ny + (ny + 1) * nx;
        c.G = G;
        c.g_fixed = g_fixed;
        c.n_lslots = (int)std::max<size_t>(1, plan->lsurf_glob.size());
        c.ptr = d_surf_ptr.p;
        c.idx = d_surf_idx.p;
        c.csum = d_csum.p;
        c.ssum = d_ssum.p;
        c.area = d_surf_area.p;
        c.cur = d_cur.p;
        c.sfl = d_sfl.p;
        return c;
    }

    RedArgs red_args() const
    {
        RedArgs a{};
        a.F = F;
        a.G = G;
        a.fsr_mat = d_fsr_mat.p;
        a.mat_nf = d_mat_nf.p;
        a.vol = d_vol.p;
        a.phi = d_phi.p;
        a.phi_old = d_phi_old.p;
        a.fs = d_fs.p;
        a.fs_prev = d_fs_prev.p;
        a.scal = d_scal.p;
        return a;
    }

    void reduce(int mode, int update_k);


    void build_source(int g_fixed, bool jacobi);


    // The sweep-kernel launches of one inner sweep of group g (g < 0: all
    // groups) on `stream`: one launch over every band (Jacobi update, the
    // caller flips the parity), or the octant-1 bands, then the octant-2
    // bands (Gauss-Seidel update).
    void launch_sweeps(int g, bool with_current);

    // one inner iteration of sweep1g for group g (g < 0: all groups)
    void sweep_once(int g, bool with_current);


    // MoCSweeper::sweep (moc_sweeper.cpp:189-225)
    void mocsweep(int g, bool with_current)
    {
        for (int inner = 0; inner < opt.n_inner; inner++)
            sweep_once(g, with_current && inner == opt.n_inner - 1);
    }

    void copy_flux_to_old();


    // FixedSourceSolver::step (fixed_source_solver.cpp:102-117)
    void fss_step(bool with_current)
    {
        copy_flux_to_old();
        if (opt.energy_mode == 1) {
            build_source(-1, true);
            mocsweep(-1, with_current);
        } else {
            for (int g = 0; g < G; g++) {
                build_source(g, false);
                mocsweep(g, with_current);
            }
        }
    }

    void sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

    CmfdView cmfd_view()
    {
        CmfdView v{};
        v.F = F;
        v.nc = nx * ny;
        v.G = G;
        v.npin = npin;
        v.n_mat = n_mat;
        v.fsr_mat = d_fsr_mat.p;
        v.pin_off = d_pin_off.p;
        v.pin_cell = d_pin_cell.p;
        v.area = d_area.p;
        v.vol = d_vol.p;
        v.pin_area = d_pin_area.p;
        v.mat_tr = d_mat_tr.p;
        v.mat_nf = d_mat_nf.p;
        v.mat_f = d_mat_f.p;
        v.mat_ch = d_mat_ch.p;
        v.mat_sc = d_mat_sc.p;
        v.phi = d_phi.p;
        v.cur = d_cur.p;
        v.sfl = d_sfl.p;
        v.k = d_scal.p + S_K;
        v.has_radial = has_radial ? 1 : 0;
        v.neg_fixup = opt.cmfd_negative_fixup ? 1 : 0;
        return v;
    }

    // EigenSolver::do_cmfd with k on the device; raises on a non-positive
    // coarse flux (set_pin_flux_1g throws there in the reference)
    int cmfd_step_sync(int *iterations)
    {
        cmfd->enqueue_step(cmfd_view(), stream);
        int bad = 0;
        cmfd->read_info(stream, iterations, &bad, &cmfd_bicgstab);
        if (bad) {
            err = "Negative or zero input flux in the CMFD rescale";
            return MOCGPU_ERR_NONPOS_FLUX;
        }
        return 0;
    }
};

