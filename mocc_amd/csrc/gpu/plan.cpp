#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace mocc_amd {
namespace gpu {

namespace {
constexpr double PI = 3.1415926535897932;
enum { EAST = 0, NORTH = 1, WEST = 2, SOUTH = 3, TOP = 4, BOTTOM = 5, INVALID = 7 };

int reflect(int iang, int normal, int nd)
{
    static const int refl[3][8] = {{1, 0, 3, 2, 5, 4, 7, 6}, {3, 2, 1, 0, 7, 6, 5, 4}, {4, 5, 6, 7, 0, 1, 2, 3}};
    int ioct = iang / nd;
    return iang + (refl[normal][ioct] - ioct) * nd;
}

// The azimuth an exit of azimuth a enters through a reflective face
// (either normal; both give the same one).
int mirror_azimuth(int a, int np, int nd)
{
    const int i = a * np;
    const int x = reflect(i, 0, nd) % (2 * nd), y = reflect(i, 1, nd) % (2 * nd);
    if (x != y) throw std::runtime_error("x and y reflections of an azimuth differ");
    return x / np;
}

// Mesh::coarse_surf / coarse_neighbor / surface_normal for one plane
struct Coarse {
    int nx, ny;
    int surf(int cell, int s) const
    {
        int ix = cell % nx, iy = cell / nx;
        switch (s) {
        case EAST: return (nx + 1) * iy + ix + 1 + nx * ny;
        case WEST: return (nx + 1) * iy + ix + nx * ny;
        case NORTH: return nx * ny + (nx + 1) * ny + (ny + 1) * ix + iy + 1;
        case SOUTH: return (nx + 1) * ny + (ny + 1) * ix + iy + nx * ny;
        }
        throw std::runtime_error("coarse crossing on a z-normal surface");
    }
    int neighbor(int cell, int s) const
    {
        int ix = cell % nx, iy = cell / nx;
        switch (s) {
        case NORTH: return (iy < ny - 1) ? cell + nx : -1;
        case SOUTH: return (iy > 0) ? cell - nx : -1;
        case EAST: return (ix < nx - 1) ? cell + 1 : -1;
        case WEST: return (ix > 0) ? cell - 1 : -1;
        }
        return -5;
    }
    int normal(int s) const
    {
        int nsurfz = nx * ny + (nx + 1) * ny + (ny + 1) * nx;
        int r = s % nsurfz;
        if (r < nx * ny) return 2;
        if (r < nx * ny + (nx + 1) * ny) return 0;
        return 1;
    }
};

struct Event {
    int step, surf, normal;
};

// moc::Current::post_ray (moc_current_worker.hpp:202-264) flattened: the
// (segment position, surface, normal) of every tally, in the order the
// direction meets them.
void ray_events(const mocgpu_problem &p, const Coarse &cm, int r, int nseg, std::vector<Event> &fw,
                std::vector<Event> &bw)
{
    int cell_fw = p.ray_cm_cell[2 * r], cell_bw = p.ray_cm_cell[2 * r + 1];
    int surf_fw = p.ray_cm_surf[2 * r], surf_bw = p.ray_cm_surf[2 * r + 1];
    int iseg_fw = 0, iseg_bw = nseg;
    fw.push_back({iseg_fw, surf_fw, cm.normal(surf_fw)});
    bw.push_back({nseg - iseg_bw, surf_bw, cm.normal(surf_bw)});
    for (int c = p.ray_cm_off[r]; c < p.ray_cm_off[r + 1]; c++) {
        const int *crd = &p.cm_data[4 * c];
        if (crd[0] != INVALID) {
            iseg_fw += crd[2];
            int s = cm.surf(cell_fw, crd[0]);
            fw.push_back({iseg_fw, s, (crd[0] == EAST || crd[0] == WEST) ? 0 : 1});
        }
        if (crd[1] != INVALID) {
            iseg_bw -= crd[3];
            int s = cm.surf(cell_bw, crd[1]);
            bw.push_back({nseg - iseg_bw, s, (crd[1] == EAST || crd[1] == WEST) ? 0 : 1});
        }
        cell_fw = cm.neighbor(cell_fw, crd[0]);
        cell_bw = cm.neighbor(cell_bw, crd[1]);
    }
}

// Per-thread scratch of the band builder: FSR and surface marks stamped with
// a per-thread band counter, per-FSR and per-surface contribution counts
// (cleared again after each band) and the current band's event lists.
struct BandScratch {
    std::vector<int> stamp, g2l, sstamp, s2l;
    std::vector<int32_t> fsr_cnt, surf_cnt;
    std::vector<std::vector<Event>> evf, evb;
    int band_id = 0;
    BandScratch(int F, int n_surf)
        : stamp(F, -1), g2l(F, -1), sstamp(n_surf, -1), s2l(n_surf, -1), fsr_cnt(F, 0), surf_cnt(n_surf, 0)
    {
    }
};

// The bands of one azimuth. Their offsets (loc_base, ray_base, seg_base,
// lsurf_base and the event offsets) count from this azimuth's own arrays
// until the azimuths are joined in order.
struct AziBands {
    std::vector<BandInfo> bands;
    std::vector<int32_t> loc_fsr, ray_global, lsurf_glob, ev_off;
    std::vector<uint16_t> seg_loc;
    std::vector<float> seg_len;
    std::vector<RayInfo> rays;
    std::vector<uint32_t> ev;
};

// Bands of azimuth a: its rays in spatial order, grown into bands whose
// touched FSRs fit the band-local index space, each with its chunked forward
// and backward streams, ray records and coarse-surface events.
void build_azimuth(const mocgpu_problem &p, const Coarse &cm, int a, int max_rays, BandScratch &sc, AziBands &pt)
{
    const int NP = p.n_polar, n_az_oct = p.n_azi / 2;
    const double hx = p.x_vec[p.nx], hy = p.y_vec[p.ny];
    const int oct = a / n_az_oct;  // 0: octant 1, 1: octant 2
    const int r0 = p.azi_ray_off[a], r1 = p.azi_ray_off[a + 1];
    const int Ny = p.azi_ny[a], Nx = p.azi_nx[a];
    const double alpha = p.ang_alpha[a * NP];
    const double sx = hx / Nx, sy = hy / Ny;
    // spatial order: distance of the entry point along the normal
    std::vector<std::pair<double, int>> order;
    for (int r = r0; r < r1; r++) {
        int b0 = p.ray_bc[2 * r];
        double x, y;
        if (b0 < Ny) {
            x = oct == 0 ? 0.0 : hx;
            y = (0.5 + b0) * sy;
        } else {
            x = (0.5 + (b0 - Ny)) * sx;
            y = 0.0;
        }
        order.push_back({-std::sin(alpha) * x + std::cos(alpha) * y, r});
    }
    std::sort(order.begin(), order.end());
    size_t i = 0;
    while (i < order.size()) {
        // grow a band
        sc.evf.clear();
        sc.evb.clear();
        std::vector<int> brays;
        std::vector<int> locs;
        int n_local = 0;
        while (i < order.size() && (int)brays.size() < max_rays) {
            int r = order[i].second;
            int new_count = 0;
            for (int64_t s = p.ray_seg_off[r]; s < p.ray_seg_off[r + 1]; s++) {
                int f = p.seg_fsr[s];
                if (sc.stamp[f] != sc.band_id && sc.g2l[f] != -2) {
                    sc.g2l[f] = -2;  // counted for this ray
                    new_count++;
                }
            }
            for (int64_t s = p.ray_seg_off[r]; s < p.ray_seg_off[r + 1]; s++) {
                int f = p.seg_fsr[s];
                if (sc.g2l[f] == -2) sc.g2l[f] = -1;
            }
            if (!brays.empty() && n_local + new_count > kMaxLocal) break;
            if (n_local + new_count > kMaxLocal)
                throw std::runtime_error("a single ray touches more FSRs than a band holds");
            for (int64_t s = p.ray_seg_off[r]; s < p.ray_seg_off[r + 1]; s++) {
                int f = p.seg_fsr[s];
                if (sc.stamp[f] != sc.band_id) {
                    sc.stamp[f] = sc.band_id;
                    locs.push_back(f);
                    n_local++;
                }
            }
            brays.push_back(r);
            i++;
        }
        // the band's rays longest first (ties in spatial order): with the
        // length-split lane map each wave takes 32 consecutive rays in both
        // directions, so a wave of short rays ends early
        std::stable_sort(brays.begin(), brays.end(), [&p](int x, int y) {
            return p.ray_seg_off[x + 1] - p.ray_seg_off[x] > p.ray_seg_off[y + 1] - p.ray_seg_off[y];
        });
        std::sort(locs.begin(), locs.end());
        for (size_t l = 0; l < locs.size(); l++) sc.g2l[locs[l]] = (int)l;
        BandInfo b{};
        b.azi = a;
        b.n_rays = (int)brays.size();
        b.n_local = n_local;
        b.loc_base = (int32_t)pt.loc_fsr.size();
        b.ray_base = (int32_t)pt.rays.size();
        b.nmax = 0;
        for (int r : brays) b.nmax = std::max(b.nmax, (int)(p.ray_seg_off[r + 1] - p.ray_seg_off[r]));
        if (b.nmax > 65535) throw std::runtime_error("ray longer than 65535 segments");
        for (int f : locs) pt.loc_fsr.push_back(f);
        // forward and backward streams, chunked layout
        const int nchunk = (b.nmax + kChunk - 1) / kChunk;
        for (int d = 0; d < 2; d++) {
            b.seg_base[d] = (int64_t)pt.seg_loc.size();
            for (int c = 0; c < nchunk; c++)
                for (int j = 0; j < b.n_rays; j++) {
                    int r = brays[j];
                    int64_t s0 = p.ray_seg_off[r];
                    int n = (int)(p.ray_seg_off[r + 1] - s0);
                    for (int u = 0; u < kChunk; u++) {
                        int t = c * kChunk + u;
                        if (t < n) {
                            int k = d == 0 ? t : n - 1 - t;
                            pt.seg_loc.push_back((uint16_t)(sc.g2l[p.seg_fsr[s0 + k]] * kSlotBytes));
                            pt.seg_len.push_back((float)p.seg_len[s0 + k]);
                        } else {
                            // padding: a zero-length step into the lane's own dummy slot
                            pt.seg_loc.push_back((uint16_t)((n_local + j) * kSlotBytes));
                            pt.seg_len.push_back(0.0f);
                        }
                    }
                }
        }
        // contributions per local FSR: every segment, both directions
        b.max_contrib = 1;
        for (int r : brays)
            for (int64_t s = p.ray_seg_off[r]; s < p.ray_seg_off[r + 1]; s++) sc.fsr_cnt[p.seg_fsr[s]] += 2;
        for (int f : locs) {
            b.max_contrib = std::max(b.max_contrib, sc.fsr_cnt[f]);
            sc.fsr_cnt[f] = 0;
        }
        // ray records and coarse events
        b.lsurf_base = (int32_t)pt.lsurf_glob.size();
        std::vector<int> lsurfs;
        for (int j = 0; j < b.n_rays; j++) {
            int r = brays[j];
            int n = (int)(p.ray_seg_off[r + 1] - p.ray_seg_off[r]);
            RayInfo ri{};
            ri.bc0 = p.ray_bc[2 * r];
            ri.bc1 = p.ray_bc[2 * r + 1];
            ri.nseg = n;
            int fw_y = ri.bc1 >= Ny ? 1 : 0;
            int bw_y = ri.bc0 >= Ny ? 1 : 0;
            // exit faces: octant 1 fw leaves E/N, octant 2 fw leaves W/N,
            // their reverses leave W/S and E/S
            int fw_face = fw_y ? NORTH : (oct == 0 ? EAST : WEST);
            int bw_face = bw_y ? SOUTH : (oct == 0 ? WEST : EAST);
            ri.flags = fw_y | (bw_y << 1) | ((p.bc[fw_face] == MOCGPU_BC_VACUUM) << 2) |
                       ((p.bc[bw_face] == MOCGPU_BC_VACUUM) << 3);
            pt.rays.push_back(ri);
            pt.ray_global.push_back(r);
            std::vector<Event> ef, eb;
            ray_events(p, cm, r, n, ef, eb);
            for (int d = 0; d < 2; d++) {
                for (const Event &e : (d == 0 ? ef : eb)) {
                    if (sc.sstamp[e.surf] != sc.band_id) {
                        sc.sstamp[e.surf] = sc.band_id;
                        lsurfs.push_back(e.surf);
                    }
                    sc.surf_cnt[e.surf]++;
                }
            }
            sc.evf.push_back(std::move(ef));
            sc.evb.push_back(std::move(eb));
        }
        b.max_surf_contrib = 1;
        for (int s : lsurfs) {
            b.max_surf_contrib = std::max(b.max_surf_contrib, sc.surf_cnt[s]);
            sc.surf_cnt[s] = 0;
        }
        std::sort(lsurfs.begin(), lsurfs.end());
        if (lsurfs.size() > (size_t)kMaxLSurf) throw std::runtime_error("too many coarse surfaces in one band");
        for (size_t l = 0; l < lsurfs.size(); l++) sc.s2l[lsurfs[l]] = (int)l;
        b.n_lsurf = (int)lsurfs.size();
        for (int s : lsurfs) pt.lsurf_glob.push_back(s);
        for (int j = 0; j < b.n_rays; j++) {
            const size_t gi = (size_t)j;
            for (int d = 0; d < 2; d++) {
                for (const Event &e : (d == 0 ? sc.evf[gi] : sc.evb[gi])) {
                    if (e.step > 65535) throw std::runtime_error("event step out of range");
                    pt.ev.push_back((uint32_t)e.step | ((uint32_t)sc.s2l[e.surf] << 16) |
                                    ((uint32_t)e.normal << 31));
                }
                pt.ev_off.push_back((int32_t)pt.ev.size());
            }
        }
        pt.bands.push_back(b);
        sc.band_id++;
    }
}

// Threads for the band builder: MOCC_HOST_THREADS, else OMP_NUM_THREADS, else
// the hardware's count (at most 64); one thread for problems under a million
// segments, where starting threads costs more than it saves.
int plan_threads(int64_t n_seg, int n_azi)
{
    if (n_seg < 1000000) return 1;
    int n = 0;
    for (const char *name : {"MOCC_HOST_THREADS", "OMP_NUM_THREADS"}) {
        const char *v = std::getenv(name);
        if (v && std::atoi(v) > 0) {
            n = std::atoi(v);
            break;
        }
    }
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min({n, 64, n_azi}));
}
}  // namespace

SweepPlan build_plan(const mocgpu_problem &p, int max_rays)
{
    if (max_rays <= 0 || max_rays > kMaxRaysPerBand) max_rays = kMaxRaysPerBand;
    SweepPlan pl;
    pl.F = p.n_fsr;
    pl.G = p.n_group;
    pl.NP = p.n_polar;
    pl.A = p.n_azi;
    pl.nd = p.n_azi * p.n_polar / 2;
    if (p.n_azi % 2 != 0 || pl.nd <= 0) throw std::runtime_error("n_azi must be even and positive");
    if (p.n_polar < 1 || p.n_polar > 4) throw std::runtime_error("1 to 4 polar angles per azimuth are supported");
    for (int s = 0; s < 4; s++)
        if (p.bc[s] != MOCGPU_BC_VACUUM && p.bc[s] != MOCGPU_BC_REFLECT)
            throw std::runtime_error("only vacuum and reflective radial boundaries are supported");
    const int nd = pl.nd, NP = pl.NP, A = pl.A;

    // boundary layout (moc_sweeper.cpp:38-52, boundary_condition.cpp:33-78)
    pl.ang_off.assign(4 * nd, 0);
    {
        int acc = 0;
        for (int iang = 0; iang < 4 * nd; iang++) {
            int i1 = iang < 2 * nd ? iang : iang - 2 * nd;
            int a = i1 / NP;
            pl.ang_off[iang] = acc;
            acc += p.azi_ny[a] + p.azi_nx[a];
        }
        pl.bcpg = acc;
    }
    pl.azi.resize(A);
    for (int a = 0; a < A; a++) {
        int i0 = a * NP;
        AziInfo &z = pl.azi[a];
        z.in_off[0] = pl.ang_off[i0];
        z.in_off[1] = pl.ang_off[i0 + 2 * nd];
        for (int n = 0; n < 2; n++) {
            z.out_off[0][n] = pl.ang_off[reflect(i0, n, nd)];
            z.out_off[1][n] = pl.ang_off[reflect(i0 + 2 * nd, n, nd)];
        }
        z.stride = p.azi_nx[a] + p.azi_ny[a];
        z.pad = 0;
    }
    pl.ang_rsin.resize(A * NP);
    pl.ang_wt.resize(A * NP);
    pl.ang_cw.resize(A * NP * 2);
    pl.ang_fw.resize(A * NP * 2);
    for (int a = 0; a < A; a++) {
        for (int pp = 0; pp < NP; pp++) {
            int iang = a * NP + pp;
            double spacing = p.azi_spacing[a];
            // sweep1g: wt_v_st = weight * spacing * height * sin(theta) * PI
            double wt = p.ang_weight[iang] * spacing * p.hz * p.ang_sin_theta[iang] * PI;
            pl.ang_rsin[iang] = (float)p.ang_rsin_theta[iang];
            pl.ang_wt[iang] = (float)wt;
            // moc::Current::set_angle
            double w = p.ang_weight[iang] * PI;
            double al = p.ang_alpha[iang];
            double cw0 = w * p.ang_ox[iang] * spacing / std::abs(std::cos(al)) * p.hz;
            double cw1 = w * p.ang_oy[iang] * spacing / std::abs(std::sin(al)) * p.hz;
            double fw0 = w * spacing / std::abs(std::cos(al)) * p.hz;
            double fw1 = w * spacing / std::abs(std::sin(al)) * p.hz;
            pl.ang_cw[2 * iang] = (float)cw0;
            pl.ang_cw[2 * iang + 1] = (float)cw1;
            pl.ang_fw[2 * iang] = (float)fw0;
            pl.ang_fw[2 * iang + 1] = (float)fw1;
        }
    }

    Coarse cm{p.nx, p.ny};
    pl.n_surf = (p.nx + 1) * p.ny + (p.ny + 1) * p.nx + 2 * p.nx * p.ny;
    pl.surf_area.assign(pl.n_surf, 0.0);
    for (int s = 0; s < pl.n_surf; s++) {
        int nsp = p.nx * p.ny + (p.nx + 1) * p.ny + (p.ny + 1) * p.nx;
        if (s >= nsp) continue;
        int n = cm.normal(s);
        if (n == 0) {
            int iy = (s - p.nx * p.ny) / (p.nx + 1);
            pl.surf_area[s] = (p.y_vec[iy + 1] - p.y_vec[iy]) * p.hz;
        } else if (n == 1) {
            int r = s - p.nx * p.ny - (p.nx + 1) * p.ny;
            int ix = r / (p.ny + 1);
            pl.surf_area[s] = (p.x_vec[ix + 1] - p.x_vec[ix]) * p.hz;
        }
    }

    // The azimuths' bands are built independently (on several threads for a
    // large problem) and joined in azimuth order, so the plan does not depend
    // on the thread count. An error is reported for the lowest azimuth that
    // has one, as a sequential build would.
    std::vector<AziBands> parts(A);
    {
        const int nt = plan_threads(p.n_seg, A);
        std::vector<std::exception_ptr> errs(A);
        std::atomic<int> next{0};
        auto work = [&]() {
            BandScratch sc(pl.F, pl.n_surf);
            for (int a = next++; a < A; a = next++) {
                try {
                    build_azimuth(p, cm, a, max_rays, sc, parts[a]);
                } catch (...) {
                    errs[a] = std::current_exception();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; t++) pool.emplace_back(work);
        work();
        for (auto &t : pool) t.join();
        for (auto &e : errs)
            if (e) std::rethrow_exception(e);
    }
    {
        size_t n_seg = 0, n_loc = 0, n_ray = 0, n_ls = 0, n_ev = 0, n_b = 0;
        for (const AziBands &pt : parts) {
            n_seg += pt.seg_loc.size();
            n_loc += pt.loc_fsr.size();
            n_ray += pt.rays.size();
            n_ls += pt.lsurf_glob.size();
            n_ev += pt.ev.size();
            n_b += pt.bands.size();
        }
        pl.seg_loc.reserve(n_seg);
        pl.seg_len.reserve(n_seg);
        pl.loc_fsr.reserve(n_loc);
        pl.rays.reserve(n_ray);
        pl.ray_global.reserve(n_ray);
        pl.lsurf_glob.reserve(n_ls);
        pl.ev.reserve(n_ev);
        pl.ev_off.reserve(2 * n_ray + 1);
        pl.bands.reserve(n_b);
    }
    pl.ev_off.push_back(0);
    for (AziBands &pt : parts) {
        const int32_t loc0 = (int32_t)pl.loc_fsr.size(), ray0 = (int32_t)pl.rays.size();
        const int32_t ls0 = (int32_t)pl.lsurf_glob.size(), ev0 = (int32_t)pl.ev.size();
        const int64_t seg0 = (int64_t)pl.seg_loc.size();
        for (BandInfo b : pt.bands) {
            b.loc_base += loc0;
            b.ray_base += ray0;
            b.lsurf_base += ls0;
            b.seg_base[0] += seg0;
            b.seg_base[1] += seg0;
            pl.bands.push_back(b);
        }
        pl.seg_loc.insert(pl.seg_loc.end(), pt.seg_loc.begin(), pt.seg_loc.end());
        pl.seg_len.insert(pl.seg_len.end(), pt.seg_len.begin(), pt.seg_len.end());
        pl.loc_fsr.insert(pl.loc_fsr.end(), pt.loc_fsr.begin(), pt.loc_fsr.end());
        pl.rays.insert(pl.rays.end(), pt.rays.begin(), pt.rays.end());
        pl.ray_global.insert(pl.ray_global.end(), pt.ray_global.begin(), pt.ray_global.end());
        pl.lsurf_glob.insert(pl.lsurf_glob.end(), pt.lsurf_glob.begin(), pt.lsurf_glob.end());
        pl.ev.insert(pl.ev.end(), pt.ev.begin(), pt.ev.end());
        for (int32_t e : pt.ev_off) pl.ev_off.push_back(e + ev0);
        pt = AziBands();  // release the azimuth's arrays
    }
    pl.n_seg_padded = (int64_t)pl.seg_loc.size();
    for (const BandInfo &b : pl.bands) {
        pl.max_local = std::max(pl.max_local, b.n_local);
        pl.max_band_rays = std::max(pl.max_band_rays, b.n_rays);
        pl.max_lsurf = std::max(pl.max_lsurf, b.n_lsurf);
    }

    // FSR -> partial slots, surface -> current slots
    {
        std::vector<int32_t> cnt(pl.F + 1, 0);
        for (int32_t f : pl.loc_fsr) cnt[f + 1]++;
        for (int f = 0; f < pl.F; f++) cnt[f + 1] += cnt[f];
        pl.fsr_slot_ptr = cnt;
        pl.fsr_slot_idx.assign(pl.loc_fsr.size(), 0);
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        for (size_t s = 0; s < pl.loc_fsr.size(); s++) pl.fsr_slot_idx[fill[pl.loc_fsr[s]]++] = (int32_t)s;
        pl.slot_pos.assign(pl.loc_fsr.size(), 0);
        for (size_t q = 0; q < pl.fsr_slot_idx.size(); q++) pl.slot_pos[pl.fsr_slot_idx[q]] = (int32_t)q;
        for (int f = 0; f < pl.F; f++)
            if (pl.fsr_slot_ptr[f + 1] == pl.fsr_slot_ptr[f])
                throw std::runtime_error("No rays passed through FSR " + std::to_string(f) +
                                         "; try a finer ray spacing");
    }
    {
        std::vector<int32_t> cnt(pl.n_surf + 1, 0);
        for (int32_t s : pl.lsurf_glob) cnt[s + 1]++;
        for (int s = 0; s < pl.n_surf; s++) cnt[s + 1] += cnt[s];
        pl.surf_slot_ptr = cnt;
        pl.surf_slot_idx.assign(pl.lsurf_glob.size(), 0);
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        for (size_t s = 0; s < pl.lsurf_glob.size(); s++) pl.surf_slot_idx[fill[pl.lsurf_glob[s]]++] = (int32_t)s;
    }
    // longest bands first
    pl.band_order.resize(pl.bands.size());
    std::iota(pl.band_order.begin(), pl.band_order.end(), 0);
    std::stable_sort(pl.band_order.begin(), pl.band_order.end(), [&](int x, int y) {
        return (int64_t)pl.bands[x].nmax > (int64_t)pl.bands[y].nmax;
    });
    // Gauss-Seidel order in angle: every band of an octant-1 azimuth, then
    // every band of an octant-2 azimuth (each longest-first)
    pl.band_order_gs.clear();
    for (int oct = 0; oct < 2; oct++) {
        for (int b : pl.band_order)
            if ((pl.bands[b].azi >= pl.A / 2) == (oct == 1)) pl.band_order_gs.push_back(b);
        if (oct == 0) pl.n_bands_oct1 = (int32_t)pl.band_order_gs.size();
    }
    return pl;
}

namespace {
// Invariants of a plan against the problem it was built from (host only).
std::string check_plan(const mocgpu_problem &p, const SweepPlan &pl, int band_rays)
{
    const int width = (band_rays <= 0 || band_rays > kMaxRaysPerBand) ? kMaxRaysPerBand : band_rays;
    auto fail = [](const std::string &m) { throw std::runtime_error("plan check: " + m); };
    std::vector<int> seen(p.n_ray, 0);
    int64_t n_seg = 0;
    for (size_t bi = 0; bi < pl.bands.size(); bi++) {
        const BandInfo &B = pl.bands[bi];
        if (B.n_rays < 1 || B.n_rays > width) fail("band width out of range");
        if (B.n_local < 1 || B.n_local > kMaxLocal) fail("band-local FSR count out of range");
        const int nchunk = (B.nmax + kChunk - 1) / kChunk;
        for (int j = 0; j < B.n_rays; j++) {
            const int r = pl.ray_global[B.ray_base + j];
            if (r < 0 || r >= p.n_ray) fail("ray index");
            seen[r]++;
            if (r < p.azi_ray_off[B.azi] || r >= p.azi_ray_off[B.azi + 1]) fail("ray outside its band's azimuth");
            const RayInfo &ri = pl.rays[B.ray_base + j];
            const int64_t s0 = p.ray_seg_off[r];
            const int n = (int)(p.ray_seg_off[r + 1] - s0);
            if (ri.nseg != n || n > B.nmax) fail("segment count");
            if (ri.bc0 != p.ray_bc[2 * r] || ri.bc1 != p.ray_bc[2 * r + 1]) fail("boundary indices");
            n_seg += n;
            for (int d = 0; d < 2; d++)
                for (int t = 0; t < nchunk * kChunk; t++) {
                    const size_t e = (size_t)B.seg_base[d] + ((size_t)(t / kChunk) * B.n_rays + j) * kChunk + t % kChunk;
                    if (e >= pl.seg_loc.size()) fail("stream index");
                    if (pl.seg_loc[e] % kSlotBytes != 0) fail("stream offset not a slot boundary");
                    const int loc = pl.seg_loc[e] / kSlotBytes;
                    const float len = pl.seg_len.empty() ? 0.0f : pl.seg_len[e];
                    if (t < n) {
                        const int k = d == 0 ? t : n - 1 - t;
                        if (loc >= B.n_local || pl.loc_fsr[B.loc_base + loc] != p.seg_fsr[s0 + k])
                            fail("segment FSR in the stream");
                        if (len != (float)p.seg_len[s0 + k]) fail("segment length in the stream");
                    } else if (loc != B.n_local + j || len != 0.0f) {
                        fail("padding step");
                    }
                }
            for (int d = 0; d < 2; d++) {
                const int64_t gi = 2 * ((int64_t)B.ray_base + j) + d;
                for (int q = pl.ev_off[gi]; q < pl.ev_off[gi + 1]; q++) {
                    const uint32_t ev = pl.ev[q];
                    if ((int)(ev & 0xFFFFu) > n) fail("event step");
                    if ((int)((ev >> 16) & 0x7FFFu) >= B.n_lsurf) fail("event surface");
                }
                if (pl.ev_off[gi + 1] < pl.ev_off[gi]) fail("event offsets");
            }
        }
    }
    for (int r = 0; r < p.n_ray; r++)
        if (seen[r] != 1) fail("every ray in exactly one band");
    if (n_seg != p.n_seg) fail("segment total");
    std::vector<int> slot_seen(pl.loc_fsr.size(), 0);
    for (int f = 0; f < p.n_fsr; f++)
        for (int s = pl.fsr_slot_ptr[f]; s < pl.fsr_slot_ptr[f + 1]; s++) {
            const int slot = pl.fsr_slot_idx[s];
            if (slot < 0 || slot >= (int)pl.loc_fsr.size() || pl.loc_fsr[slot] != f) fail("FSR slot list");
            if (s > pl.fsr_slot_ptr[f] && slot <= pl.fsr_slot_idx[s - 1]) fail("FSR slots not increasing");
            if (pl.slot_pos.size() != pl.loc_fsr.size() || pl.slot_pos[slot] != s) fail("slot positions");
            slot_seen[slot]++;
        }
    for (int v : slot_seen)
        if (v != 1) fail("every slot listed once");
    std::vector<int> o1(pl.bands.size(), 0), o2(pl.bands.size(), 0);
    for (size_t k = 0; k < pl.band_order.size(); k++) {
        o1[pl.band_order[k]]++;
        if (k && pl.bands[pl.band_order[k]].nmax > pl.bands[pl.band_order[k - 1]].nmax) fail("band order");
    }
    for (size_t k = 0; k < pl.band_order_gs.size(); k++) {
        const int b = pl.band_order_gs[k];
        o2[b]++;
        if ((pl.bands[b].azi >= p.n_azi / 2) != ((int)k >= pl.n_bands_oct1)) fail("Gauss-Seidel phase order");
    }
    // The two-phase Gauss-Seidel update needs every exit of an azimuth to
    // land in the incoming slots of its mirror, an azimuth of the other octant.
    const int np_ = p.n_polar, nd_ = p.n_azi * p.n_polar / 2;
    for (int az = 0; az < p.n_azi; az++) {
        const int m = mirror_azimuth(az, np_, nd_);
        if (m < 0 || m >= p.n_azi || (m >= p.n_azi / 2) == (az >= p.n_azi / 2)) fail("mirror azimuth");
        for (int n = 0; n < 2; n++)
            for (int d = 0; d < 2; d++) {
                const int out = pl.azi[az].out_off[d][n];
                if (out != pl.azi[m].in_off[0] && out != pl.azi[m].in_off[1]) fail("exit offsets miss the mirror azimuth");
            }
    }
    for (size_t b = 0; b < pl.bands.size(); b++)
        if (o1[b] != 1 || o2[b] != 1) fail("band orders are permutations");
    // FNV-1a digest of the arrays the device receives, so two builds of one
    // problem can be compared from the report
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const auto &v) {
        const unsigned char *b = (const unsigned char *)v.data();
        const size_t n = v.size() * sizeof(v[0]);
        for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
        h = (h ^ (uint64_t)n) * 1099511628211ull;
    };
    mix(pl.bands);
    mix(pl.band_order);
    mix(pl.band_order_gs);
    mix(pl.seg_loc);
    mix(pl.seg_len);
    mix(pl.loc_fsr);
    mix(pl.rays);
    mix(pl.slot_pos);
    mix(pl.fsr_slot_ptr);
    mix(pl.fsr_slot_idx);
    mix(pl.lsurf_glob);
    mix(pl.ev_off);
    mix(pl.ev);
    mix(pl.surf_slot_ptr);
    mix(pl.surf_slot_idx);
    char dig[17];
    std::snprintf(dig, sizeof dig, "%016llx", (unsigned long long)h);
    return "ok: " + std::to_string(pl.bands.size()) + " bands, " + std::to_string(pl.loc_fsr.size()) +
           " slots, " + std::to_string(pl.n_seg_padded) + " stream steps, digest " + dig;
}
}  // namespace

}  // namespace gpu
}  // namespace mocc_amd

extern "C" int mocgpu_plan_check(const mocgpu_problem *p, int32_t band_rays, char *msg, int32_t cap)
{
    std::string m;
    int rc = 0;
    try {
        const mocc_amd::gpu::SweepPlan pl = mocc_amd::gpu::build_plan(*p, band_rays);
        m = mocc_amd::gpu::check_plan(*p, pl, band_rays);
    } catch (const std::exception &e) {
        m = e.what();
        rc = MOCGPU_ERR_ARG;
    }
    if (msg && cap > 0) {
        std::strncpy(msg, m.c_str(), (size_t)cap - 1);
        msg[cap - 1] = 0;
    }
    return rc;
}
