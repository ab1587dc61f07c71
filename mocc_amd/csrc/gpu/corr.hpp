// The 2-D/3-D correction worker on the device: cmdo::CurrentCorrections
// (src/sweepers/cmdo/correction_worker.hpp:36-292, correction_worker.cpp:33-160)
// as run by MoCSweeper_2D3D::sweep (src/sweepers/cmdo/moc_sweeper_2d3d.cpp:48-99),
// and MoCSweeper::update_incoming_flux (src/sweepers/moc/moc_sweeper.cpp:255-291).
//
// The correction sweep is not the hot path (mocc runs it on the last inner
// iteration of a group when the MoC sweeper feeds a 3-D corrected
// diamond-difference Sn sweeper), so it is a separate FP64 kernel rather
// than a mode of the band sweep: after the band sweep of that inner
// iteration, one thread per (ray, direction, polar angle) sweeps its ray
// again from a copy of the same incoming angular fluxes and transport source,
// in the same boundary-update order, and records per coarse-cell visit the
// segment-integrated flux, its Sigma_t-weighted twin and the track length,
// and the angular flux at every coarse-surface crossing. Fixed-order sums over
// host-built visit and crossing lists then give each (cell, angle)'s
// correction factors and each surface's current and surface flux.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../../include/mocgpu.h"
#include "plan.hpp"

namespace mocc_amd {
namespace gpu {

// What the correction kernels read of the sweeper handle for one group.
struct CorrView {
    int F, G, g, bcpg, gs, parity;
    const float *psi;        // incoming angular flux of group g, the sweep's buffer
    const double *src, *phi; // one-group source and scalar flux [F][G]
    const int32_t *fsr_mat;
    const double *mat_tr, *mat_ss;
    const double *htr;       // homogenized Sigma_t [n_cell][G]
    double *cur, *sfl;       // CoarseData current / surface flux [n_surf][G]
};

class CorrSolver {
public:
    ~CorrSolver();
    // The rays' boundary indices, azimuths and first coarse cells with the
    // band plan's boundary offsets (what update_incoming needs); with full,
    // also the ray-major segments, coarse-cell visits and surface crossings
    // of the correction sweep.
    void setup(const mocgpu_problem &p, const SweepPlan &plan, hipStream_t s, bool full);
    bool full() const { return full_; }
    // The correction sweep of group g (CurrentCorrections worker on the last
    // inner iteration): snapshot of the inputs, the FP64 re-sweep, the
    // factors, the residuals and the coarse currents. Enqueue the snapshot
    // with before_sweep before the band sweep of that inner iteration.
    void before_sweep(const CorrView &v, hipStream_t s);
    void after_sweep(const CorrView &v, hipStream_t s);
    // CorrectionData (correction_data.hpp:45-100): alpha [G][4nd][cell][2],
    // beta [G][4nd][cell]; the residual triple of group g's last correction sweep.
    void corrections(hipStream_t s, double *alpha, double *beta) const;
    void residual(hipStream_t s, int g, double *out3) const;
    // MoCSweeper::update_incoming_flux on the sweeper's angular flux buffers.
    void update_incoming(float *psi, uint64_t parity, int G, const double *part, const double *part_old,
                         bool has_old, hipStream_t s) const;
    size_t bytes() const;

private:
    void free_all();
    int R_ = 0, A_ = 0, NP_ = 0, nd_ = 0, nx_ = 0, ny_ = 0, n_cell_ = 0, n_surf_ = 0, n_cm_ = 0, G_ = 0;
    int bcpg_ = 0;
    int64_t S_ = 0;
    bool full_ = false;
    std::vector<void *> owned_;
    // geometry
    int32_t *d_seg_fsr_ = nullptr;
    double *d_seg_len_ = nullptr;
    int64_t *d_ray_seg_off_ = nullptr;
    int32_t *d_ray_info_ = nullptr;   // [R][8]: bc0, bc1, azimuth, flags, cm_off, cm_end, cell_fw, cell_bw
    int32_t *d_ent_nseg_ = nullptr;   // [n_cm][2]: segments of the entry's run (0: direction has none)
    int32_t *d_vis_ptr_ = nullptr, *d_vis_item_ = nullptr;    // per (azimuth, cell): entry << 1 | dir
    int32_t *d_crs_ptr_ = nullptr, *d_crs_item_ = nullptr;    // per (azimuth, surface): crossings
    AziInfo *d_azi_ = nullptr;
    double *d_ang_ = nullptr;         // [A][P][8]: rsin, alpha, spacing, cw_x, cw_y, fw_x, fw_y, ox
    double *d_dx_ = nullptr, *d_dy_ = nullptr, *d_area_ = nullptr;
    // per-sweep work arrays
    double *d_q_ = nullptr, *d_bc_ = nullptr;
    double *d_ent_ = nullptr;         // [n_cm][2][P][4]: flux-volume, Sigma_t-weighted, track length, psi at the crossing
    double *d_ray_psi0_ = nullptr;    // [R][2][P]: psi at the ray's entry crossing
    double *d_res_part_ = nullptr;    // [A*P*n_cell][3]
    // results
    double *d_alpha_ = nullptr, *d_beta_ = nullptr, *d_resid_ = nullptr;
};

}  // namespace gpu
}  // namespace mocc_amd
