// Host-side sweep plan: turns the ray-major SoA problem into the band-major
// layout the gfx950 sweep kernel streams.
//
//  * A band is a run of spatially adjacent parallel rays of one azimuth whose
//    touched FSRs fit a band-local index space of at most kMaxLocal entries,
//    so one workgroup can tally them in LDS.
//  * Segments of a band are stored twice, in forward order and in backward
//    order, each in a chunked layout: step t of ray j lives at
//    seg_base[d] + ((t / kChunk) * n_rays + j) * kChunk + t % kChunk, so one
//    lane fetches kChunk steps with one 16-byte and two 16-byte loads and the
//    lanes of a wave read one contiguous block; short rays are padded with
//    zero-length steps into the dummy slot n_local + j of ray j (one per
//    lane, so padding lanes never share an LDS address).
//  * Every band-local FSR owns one partial slot; fsr_slot_ptr/idx list the
//    slots of each FSR in increasing order for the final integer reduction.
//  * Coarse-surface crossings (moc::Current) are flattened into per-ray,
//    per-direction event lists of (step, band-local surface, normal).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/mocgpu.h"

namespace mocc_amd {
namespace gpu {

constexpr int kMaxLocal = 2047;  // band-local FSRs held in LDS
constexpr int kMaxRaysPerBand = 64;
constexpr int kDummySlots = kMaxRaysPerBand;  // padding slots after a band's n_local FSRs
constexpr int kSlotBytes = 8;     // bytes per slot of the band's LDS tally and (Sigma_t, q) arrays
static_assert((kMaxLocal + kDummySlots) * kSlotBytes < 65536, "stream byte offsets must fit 16 bits");
constexpr int kChunk = 8;         // segments per lane per vector load (chunked layout)
constexpr int kMaxLSurf = 1024;   // band-local coarse surfaces held in LDS

struct BandInfo {
    int32_t azi;        // azimuth index
    int32_t n_rays;     // rays in the band
    int32_t nmax;       // longest ray (segments)
    int32_t n_local;    // band-local FSRs
    int64_t seg_base[2];// first element of the forward / backward stream
    int32_t loc_base;   // first local FSR / partial slot
    int32_t ray_base;   // first entry in ray_info
    int32_t n_lsurf;    // band-local coarse surfaces
    int32_t lsurf_base; // first local surface / current slot
    int32_t max_contrib;      // max tally contributions into one local FSR
    int32_t max_surf_contrib; // max current tallies into one local surface
};

// Ray as the kernel needs it: entry/exit boundary indices, length and the
// exit faces of both directions.
struct RayInfo {
    int32_t bc0, bc1, nseg, flags;  // flags: bit0 fw exit on a y-normal face,
                                    // bit1 bw exit on a y-normal face,
                                    // bit2 fw exit vacuum, bit3 bw exit vacuum
};

struct AziInfo {
    int32_t in_off[2];      // boundary offset of (a, p=0) entry angle, fw / bw
    int32_t out_off[2][2];  // [dir][normal] offset of the reflected angle
    int32_t stride;         // Nx + Ny (offset step between polar angles)
    int32_t pad;
};

struct SweepPlan {
    int F = 0, G = 0, NP = 0, A = 0, nd = 0;
    int n_surf = 0;
    int64_t n_seg_padded = 0;
    int max_local = 0, max_lsurf = 0;       // LDS sizing
    int max_band_rays = 0;                  // widest band (workgroup width)
    std::vector<BandInfo> bands;
    std::vector<int32_t> band_order;        // longest-first
    // Gauss-Seidel in angle: reflection couples octant-1 azimuth a only with
    // its mirror a + A/2 (boundary_condition.cpp's reflect), so mocc's
    // angle-ordered boundary update is two phases: the octant-1 bands
    // (band_order_gs[0, n_bands_oct1)), then the octant-2 bands, which read
    // the boundary fluxes phase 1 wrote.
    std::vector<int32_t> band_order_gs;
    int32_t n_bands_oct1 = 0;
    std::vector<uint16_t> seg_loc;          // interleaved band-local FSR slot x kSlotBytes (LDS byte offset)
    std::vector<float> seg_len;             // interleaved length
    std::vector<int32_t> loc_fsr;           // [total local] global FSR
    std::vector<RayInfo> rays;              // [sum n_rays]
    std::vector<int32_t> ray_global;        // [sum n_rays] problem ray of each band lane (host only)
    // Partial slots of each FSR in increasing slot order: positions
    // fsr_slot_ptr[f] .. fsr_slot_ptr[f + 1] - 1 hold slots fsr_slot_idx[..];
    // slot_pos is the inverse map (slot -> position). The band tallies are
    // stored at their positions, so each FSR's partials are contiguous.
    std::vector<int32_t> fsr_slot_ptr, fsr_slot_idx, slot_pos;
    std::vector<AziInfo> azi;
    std::vector<float> ang_rsin, ang_wt;    // [A*NP]: 1/sin(theta), w*spacing*hz*sin(theta)*pi
    std::vector<float> ang_cw, ang_fw;      // [A*NP*2]: Current weights per normal
    int bcpg = 0;                           // boundary values per group
    std::vector<int32_t> ang_off;           // [4*nd] boundary offset of each angle

    // currents
    std::vector<int32_t> lsurf_glob;        // [total local surfaces] global surface
    std::vector<int32_t> ev_off;            // [2 * total rays + 1] event list offsets (fw, bw per ray)
    std::vector<uint32_t> ev;               // packed: step (16) | local surf (15) | normal (1)
    std::vector<int32_t> surf_slot_ptr, surf_slot_idx;
    std::vector<double> surf_area;          // [n_surf] (0 for z-normal)
};

// Build the plan; throws std::runtime_error on inputs the kernels cannot take.
SweepPlan build_plan(const mocgpu_problem &p, int max_rays_per_band);


}  // namespace gpu
}  // namespace mocc_amd
