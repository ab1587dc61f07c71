// Device kernels of the gfx950 MoC sweeper. Declarations shared by the
// kernel translation unit and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.hpp"

namespace mocc_amd {
namespace gpu {

constexpr int kRays = kMaxRaysPerBand;  // lanes per direction
constexpr int kWG = 2 * kRays;          // workgroup: fw wave(s) then bw wave(s)

struct SweepArgs {
    const BandInfo *bands;
    const int32_t *order;   // band order (longest first)
    int32_t n_bands;
    int32_t g_fixed;        // >= 0: one group; < 0: all groups (item = band_rank*G + g)
    int32_t G, F, bcpg;
    int32_t lds_local;      // LDS FSR slots per workgroup (max band-local FSRs + 1)
    int32_t lds_lsurf;      // LDS coarse-surface slots (0 without currents)
    int32_t n_slots;        // partial slots per group (psum is [G][n_slots])
    int32_t n_lslots;       // current slots per group (csum/ssum are [G][n_lslots])
    uint64_t parity;        // bit g: which boundary buffer of group g is incoming
    const uint16_t *seg_loc;
    const float *seg_len;
    const int32_t *loc_fsr;
    const RayInfo *rays;
    const AziInfo *azi;
    const float *ang_rsin, *ang_wt, *ang_cw, *ang_fw;
    const float2 *exptab;
    const int32_t *fsr_mat;
    const double *mat_tr, *mat_ss;
    const double *src, *phi;
    float *psi;             // [G][2][bcpg]
    double *psum;           // [G][n_slots]
    const int32_t *ev_off;
    const uint32_t *ev;
    double *csum, *ssum;    // [G][n_lslots]
};

struct ReduceArgs {
    int32_t F, G, g_fixed, n_slots;
    const int32_t *slot_ptr, *slot_idx;
    const double *psum;
    const int32_t *fsr_mat;
    const double *mat_tr, *mat_ss, *vol;
    const double *src;
    double *phi;
};

void launch_sweep(const SweepArgs &a, int n_polar, bool with_current, int n_items, hipStream_t s);
size_t sweep_lds_bytes(const SweepArgs &a);
// Workgroups of a sweep launch: one per band (one group), or the XCD-aware
// (band, group) grid of all groups.
int sweep_grid_items(int n_bands, int G, bool all_groups);
void launch_reduce_flux(const ReduceArgs &a, hipStream_t s);

}  // namespace gpu
}  // namespace mocc_amd
