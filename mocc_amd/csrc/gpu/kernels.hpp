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
    int32_t lds_local;      // LDS FSR slots per workgroup (max band-local FSRs + kDummySlots)
    int32_t lds_lsurf;      // LDS coarse-surface slots (0 without currents)
    int32_t n_slots;        // partial slots per group
    int32_t n_lslots;       // current slots per group (csum/ssum are [G][n_lslots])
    uint64_t parity;        // bit g: which boundary buffer of group g is incoming
    const uint64_t *parity_dev;  // if set, the parity word is read from device memory instead
    int32_t gs;             // 1: Gauss-Seidel boundary update (exit fluxes go to the incoming buffer)
    int32_t interleave;     // lane map: 0 plain, 1 ray-parity interleave, 2 length split
    int32_t wg_rays;        // rays per direction of a workgroup: 64 (two waves) or 32 (one wave); the plan's widest band fits
    const uint16_t *seg_loc;
    const float *seg_len;
    const int32_t *loc_fsr;
    const RayInfo *rays;
    const AziInfo *azi;
    const float *ang_rsin, *ang_wt, *ang_cw, *ang_fw;
    const int32_t *fsr_mat;
    const double *mat_tr, *mat_ss;
    const double *src, *phi;
    float *psi;             // [G][2][bcpg]
    // Band partials, slot s of group g at psum[slot_pos[s] * psum_sp + g * psum_sg]:
    // position-major ([n_slots][G]: sp = G, sg = 1) when all groups are swept
    // together, group-major ([G][n_slots]: sp = 1, sg = n_slots) when one is
    double *psum;
    int64_t psum_sp, psum_sg;
    const int32_t *slot_pos;
    const int32_t *ev_off;
    const uint32_t *ev;
    double *csum, *ssum;    // [G][n_lslots]
};

struct ReduceArgs {
    int32_t F, G, g_fixed, n_slots;
    const int32_t *slot_ptr;  // partials of FSR f: positions s in [slot_ptr[f], slot_ptr[f + 1])
    const double *psum;       // position s of group g at psum[s * psum_sp + g * psum_sg] (as written)
    int64_t psum_sp, psum_sg;
    const int32_t *fsr_mat;
    const double *mat_tr, *mat_ss, *vol;
    const double *src;
    double *phi;
    uint64_t *parity;       // batch: boundary parity word, xor-ed with gmask after the reduction
    uint64_t gmask;
};

// Block -> (problem, first element) map of a launch over many problems; each
// problem's elements are covered by whole 256-thread blocks.
struct BlockMap {
    int32_t c, first;
};

// One persistent launch over a work queue of (case, band, group) items of many
// problems: workgroups take items from *counter until n_items is reached.
// Items of cases whose status[2 * case + 1] is non-zero (finished) or whose
// n_inner[case] <= inner are skipped.
struct QueueArgs {
    const SweepArgs *cases;
    const int4 *items;        // (case, band, group, 0)
    int32_t n_items;
    int32_t inner;
    unsigned int *counter;
    const int32_t *status;    // [2 * case + 1]: non-zero once the case has finished
    const int32_t *n_inner;   // [case]
    const int32_t *want_cur;  // [case] or null: tally currents on the case's last inner iteration
};

// Coarse-surface current reduction of one problem (surfaces [n0, n1)).
struct CurrentArgs {
    int32_t n0, n1, G, g_fixed, n_lslots;
    const int32_t *ptr, *idx;  // surface -> current slots
    const double *csum, *ssum, *area;
    double *cur, *sfl;         // [n_surf][G]
    int32_t *radial_flag;      // batch: set to 1 once currents exist
};

void launch_sweep(const SweepArgs &a, int n_polar, bool with_current, int n_items, hipStream_t s);
// Workgroups resident per CU for the queue kernel with lds bytes of LDS and
// wg_rays rays per direction (32: one-wave workgroups, else 64).
int queue_blocks_per_cu(int n_polar, size_t lds, bool cur, int wg_rays);
void launch_sweep_queue(const QueueArgs &q, int n_polar, size_t lds, int n_blocks, hipStream_t s, bool cur,
                        int wg_rays);
void launch_reduce_current(const CurrentArgs &a, size_t n_items, hipStream_t s);
void launch_reduce_current_batch(const CurrentArgs *cases, const BlockMap *bm, int n_blocks, const int32_t *status,
                                 const int32_t *n_inner, int inner, hipStream_t s);
size_t sweep_lds_bytes(const SweepArgs &a);
// Workgroups of a sweep launch: one per band (one group), or the XCD-aware
// (band, group) grid of all groups.
int sweep_grid_items(int n_bands, int G, bool all_groups);
void launch_reduce_flux(const ReduceArgs &a, hipStream_t s);
// out[i] = the sweep kernel's 1 - exp(-tau[i]) for n values.
void launch_attenuation(const float *tau, float *out, int n, hipStream_t s);

// Lanes per (FSR, group) item of the flux reduction.
int reduce_flux_lanes();
// Lanes per (surface, group) item of the current reduction.
int reduce_current_lanes();
void launch_reduce_flux_batch(const ReduceArgs *cases, const BlockMap *bm, int n_blocks, const int32_t *status,
                              const int32_t *n_inner, int inner, hipStream_t s);

}  // namespace gpu
}  // namespace mocc_amd
