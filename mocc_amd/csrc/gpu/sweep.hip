// gfx950 MoC sweep kernel: one workgroup per (ray band, group).
//
// Restates moc::MoCSweeper::sweep1g (src/sweepers/moc/moc_sweeper_kernel.inc.hpp:36-180)
// and moc::Current::post_ray (src/sweepers/moc/moc_current_worker.hpp:202-264)
// for the GPU:
//
//   * lanes 0..63 sweep the band's rays forward, lanes 64..127 backward, each
//     lane carrying all polar angles of its ray (independent recurrences);
//   * the band's cross sections and transport source q = (S + Ss*phi)/(4 pi St)
//     (source_isotropic.cpp:21-57) are staged in LDS in FP32;
//   * segments stream in chunks of kChunk steps per lane (three 16-byte loads),
//     the next chunk in flight while the current one is swept; within a chunk
//     all attenuations (independent of psi) are evaluated in registers, from
//     the reference's interpolated exponential, before the psi recurrence
//     runs;
//   * every tally contribution sum_p w_p * dpsi_p is rounded to a 64-bit
//     fixed-point integer and added into a band-local LDS tally with integer
//     atomics. Integer addition is associative, so the band's sums do not
//     depend on lane or wave timing; the band's scale 2^k is derived from a
//     bound on |dpsi| (the staged q and the entry psi) so no sum can overflow;
//   * outgoing angular fluxes are written straight into the next incoming
//     buffer of the reflected angle (BoundaryCondition::update with Jacobi
//     ordering, boundary_condition.cpp:155-191); vacuum faces stay zero.
#include "kernels.hpp"

namespace mocc_amd {
namespace gpu {

namespace {

constexpr double FPI = 4.0 * 3.1415926535897932;
constexpr int kNumXcd = 8;  // MI355X: workgroup i is dispatched to XCD i % 8

// 1 - exp(-tau) as Exponential_Linear<10000> on [-10, 0] evaluates exp(-tau)
// (src/core/exponential.hpp:56-112), from the coordinate w = 1000 tau: with
// j = floor(w), f = w - j and x = j / 1000 the interpolant is
//   H_j + f (H_{j+1} - H_j),  H_j = 1 - e^{-x},  H_{j+1} - H_j = e^{-x} c,
// c = 1 - e^{-1/1000}. e^{-x} is the hardware exp2; H_j is 1 - e^{-x} for
// j >= 250 (x >= 1/4) and x P(x) below (no cancellation near 1), P the
// degree-4 minimax fit of (1 - e^{-x}) / x on [0, 1/4] in relative error
// (2.5e-9), written in j (coefficient k times 1000^-(k+1)). Past tau = 10 the
// reference calls exp itself; the same interpolant continues there (error
// < 6e-12), and for large tau e^{-x} underflows to 0 and the result is 1.
// Relative error against the FP64 interpolant: below 3e-7 for every finite
// tau >= 0.
__device__ __forceinline__ float attenuation(float w)
{
    const float f = __builtin_amdgcn_fractf(w);
    const float j = w - f;
    const float e = __builtin_amdgcn_exp2f(j * -1.4426950408889634e-3f);
    float s = 7.510221563279629e-18f;
    s = fmaf(s, j, -4.148895293474198e-14f);
    s = fmaf(s, j, 1.6665081679821016e-10f);
    s = fmaf(s, j, -4.999994933605194e-07f);
    s = fmaf(s, j, 1.0e-3f);
    const float h = j < 250.0f ? j * s : 1.0f - e;
    return fmaf(e * 9.995001666250084e-4f, f, h);
}

// Power-of-two scale 2^k of a band's fixed-point tallies. `single` bounds
// one contribution and `count` the contributions any slot receives: every
// contribution times 2^k stays below 2^51 (the range of to_fixed) and every
// slot's sum below 2^62.
__device__ __forceinline__ int fixed_exp(double single, double count)
{
    if (!(single > 0.0)) return 0;
    int e1, e2;
    frexp(single, &e1);          // single < 2^e1
    frexp(single * count, &e2);  // sum < 2^e2
    return min(51 - e1, 62 - e2);
}

// round(v * 2^k) as a two's-complement 64-bit integer for |v * 2^k| <= 2^51,
// scale = 2^k. (double)v * 2^k is exact; adding M = 1.5 * 2^52 rounds it to
// an integer n held in the low mantissa bits, so the bit pattern of the sum
// is bits(M) + n. bits(M) = 0x4338000000000000 has a zero low word, so taking
// it off changes the high word only (three VALU instructions in all).
__device__ __forceinline__ unsigned long long to_fixed(float v, double scale)
{
    const double y = fma((double)v, scale, 6755399441055744.0);
    return (unsigned long long)__double_as_longlong(y) - 0x4338000000000000ull;
}

// One (band, group) item: the whole workgroup sweeps the band's rays in both
// directions for group g. Ends after the band's partial tallies are stored;
// the caller synchronizes before reusing smem.
template <int NP, bool CUR, int W>
__device__ __forceinline__ void sweep_band(const SweepArgs &A, int band, int g, unsigned long long *smem,
                                           unsigned int &s_bound)
{
    unsigned long long *tal = smem;
    float2 *sq = (float2 *)(smem + A.lds_local);  // (Sigma_t, q) per band-local FSR
    unsigned long long *tcur = smem + 2 * A.lds_local;
    unsigned long long *tsfl = tcur + A.lds_lsurf;

    const BandInfo B = A.bands[band];
    const int tid = threadIdx.x;
    const int G = A.G;
    // currents are tallied when the instantiation has them and this problem
    // wants them (the batch queue mixes problems with and without CMFD)
    const bool cur = CUR && A.lds_lsurf > 0;
    if (tid == 0) s_bound = 0u;
    __syncthreads();

    // stage the band's cross sections and transport source
    unsigned int lmax = 0u;
    for (int l = tid; l < B.n_local; l += 2 * W) {
        const int f = A.loc_fsr[B.loc_base + l];
        const int m = A.fsr_mat[f];
        const double tr = A.mat_tr[m * G + g];
        const double q = (A.src[(size_t)f * G + g] + A.phi[(size_t)f * G + g] * A.mat_ss[m * G + g]) *
                         (1.0 / (tr * FPI));
        const float q32 = (float)q;
        sq[l] = make_float2((float)tr, q32);
        tal[l] = 0ull;
        lmax = max(lmax, __float_as_uint(fabsf(q32)));
    }
    for (int t = tid; t < kDummySlots; t += 2 * W) {
        sq[B.n_local + t] = make_float2(0.0f, 0.0f);
        tal[B.n_local + t] = 0ull;
    }
    if (cur) {
        for (int l = tid; l < B.n_lsurf; l += 2 * W) {
            tcur[l] = 0ull;
            tsfl[l] = 0ull;
        }
    }

    // Lane map. Plain (0): one wave sweeps the band forward, the other
    // backward. Interleaved (1): each wave takes every ray, even rays in one
    // direction and odd rays in the other (the other wave the reverse), so
    // neighbouring rays enter a given FSR at different steps and fewer lanes
    // of a wave hit the same LDS tally address in the same step. Length split
    // (2): wave w takes rays 32w .. 32w + 31 in both directions, interleaved
    // the same way (lanes 0-31 even rays forward and odd rays backward, lanes
    // 32-63 the reverse); the plan orders a band's rays longest first, so the
    // second wave runs only as long as its own longest ray, and not at all in
    // a band of up to 32 rays. With one wave per band (W = 32), maps 0 and 1
    // are the plain map and map 2 its interleaved form.
    const int j = A.interleave == 2 ? (tid >> 6) * 32 + (tid & 31) : tid % W;
    const int dir = A.interleave == 2 ? (((tid >> 5) ^ tid) & 1)
                  : A.interleave ? ((j & 1) ^ (tid / W)) : (tid / W);
    const bool active = j < B.n_rays;
    const AziInfo az = A.azi[B.azi];
    // select (not index) the per-direction offsets: indexing a local array
    // with a runtime value would place it in scratch
    const int in_off = dir ? az.in_off[1] : az.in_off[0];
    const int out_x = dir ? az.out_off[1][0] : az.out_off[0][0];
    const int out_y = dir ? az.out_off[1][1] : az.out_off[0][1];
    const size_t gbase = (size_t)g * 2 * A.bcpg;
    const uint64_t parity = A.parity_dev ? *A.parity_dev : A.parity;
    const int par = (int)((parity >> g) & 1ull);
    // Jacobi: exit fluxes go to the other buffer, read by the next sweep.
    // Gauss-Seidel: they go straight to the incoming buffer; a phase only
    // writes the incoming slots of the other phase's angles.
    const float *psi_in = A.psi + gbase + (size_t)par * A.bcpg;
    float *psi_out = A.psi + gbase + (size_t)(A.gs ? par : 1 - par) * A.bcpg;

    // rk[p] = 1000 / sin(theta_p): tau * 1000 is the interpolation coordinate
    float psi[NP], rk[NP], wt[NP];
    float cw[NP][2], fw[NP][2];
    RayInfo r{0, 0, 0, 0};
    float wsum = 0.0f, cwsum = 0.0f;
#pragma unroll
    for (int p = 0; p < NP; p++) {
        rk[p] = A.ang_rsin[B.azi * NP + p] * 1000.0f;
        wt[p] = A.ang_wt[B.azi * NP + p];
        wsum += wt[p];
        psi[p] = 0.0f;
        if (CUR) {
#pragma unroll
            for (int n = 0; n < 2; n++) {
                cw[p][n] = A.ang_cw[(B.azi * NP + p) * 2 + n];
                fw[p][n] = A.ang_fw[(B.azi * NP + p) * 2 + n];
            }
            cwsum += fmaxf(fmaxf(fabsf(cw[p][0]), fabsf(cw[p][1])), fmaxf(fw[p][0], fw[p][1]));
        }
    }
    if (active) {
        r = A.rays[B.ray_base + j];
        const int entry = dir ? r.bc1 : r.bc0;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            psi[p] = psi_in[in_off + p * az.stride + entry];
            lmax = max(lmax, __float_as_uint(fabsf(psi[p])));
        }
    }
    atomicMax(&s_bound, lmax);
    __syncthreads();

    // |dpsi| <= 2 * max(|psi_in|, |q|) since psi relaxes towards q; the
    // per-contribution bound times the contribution count bounds every sum.
    const double psi_bound = (double)__uint_as_float(s_bound);
    const int kx = fixed_exp(2.0 * psi_bound * (double)wsum * 1.0001, (double)B.max_contrib);
    const double sx = ldexp(1.0, kx);
    int kc = 0, ep = 0, ee = 0;
    double sc = 1.0;
    // The lane's next two coarse-surface events, held in registers: neither a
    // step without an event nor the test for a second event at the same step
    // waits on a load. The event after them is loaded when one is consumed
    // and is needed only at the next consumption. nstep = 0x10000 is past
    // every step.
    uint32_t nev = 0u, nev2 = 0u;
    int nstep = 0x10000;
    if (cur) {
        kc = fixed_exp(psi_bound * (double)cwsum * 1.0001, (double)B.max_surf_contrib);
        sc = ldexp(1.0, kc);
        if (active) {
            ep = A.ev_off[2 * (B.ray_base + j) + dir];
            ee = A.ev_off[2 * (B.ray_base + j) + dir + 1];
            if (ep < ee) {
                nev = A.ev[ep];
                nstep = (int)(nev & 0xFFFFu);
            }
            if (ep + 1 < ee) nev2 = A.ev[ep + 1];
        }
    }
    auto next_event = [&]() {
        ep++;
        nev = nev2;
        nstep = ep < ee ? (int)(nev & 0xFFFFu) : 0x10000;
        if (ep + 1 < ee) nev2 = A.ev[ep + 1];
    };
    const float csign = dir ? -1.0f : 1.0f;
    const int n = r.nseg;
    const int nr = B.n_rays;

    auto tally_current = [&](uint32_t evv, const float(&pv)[NP]) {
        const int ls = (int)((evv >> 16) & 0x7FFFu);
        const int nn = (int)(evv >> 31);
        float cv = 0.0f, sv = 0.0f;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            cv = fmaf(pv[p], nn ? cw[p][1] : cw[p][0], cv);
            sv = fmaf(pv[p], nn ? fw[p][1] : fw[p][0], sv);
        }
        atomicAdd(&tcur[ls], to_fixed(csign * cv, sc));
        atomicAdd(&tsfl[ls], to_fixed(sv, sc));
    };

    // chunked segment stream of this lane's ray and direction
    const int64_t sbase = dir ? B.seg_base[1] : B.seg_base[0];
    const uint4 *__restrict__ locp = (const uint4 *)(A.seg_loc + sbase);
    const float4 *__restrict__ lenp = (const float4 *)(A.seg_len + sbase);
    const int nchunk = (n + kChunk - 1) / kChunk;
    uint4 cl = make_uint4(0, 0, 0, 0);
    float4 ca = make_float4(0, 0, 0, 0), cb = ca;
    // lane's current chunk; chunk c of ray j is row c of nr rays
    const uint4 *__restrict__ lp = locp + j;
    const float4 *__restrict__ ap = lenp + 2 * j;
    if (active && nchunk > 0) {
        cl = lp[0];
        ca = ap[0];
        cb = ap[1];
    }
    for (int c = 0; c < nchunk; c++) {
        unsigned int loc[kChunk];
        float len[kChunk];
        loc[0] = cl.x & 0xFFFFu;
        loc[1] = cl.x >> 16;
        loc[2] = cl.y & 0xFFFFu;
        loc[3] = cl.y >> 16;
        loc[4] = cl.z & 0xFFFFu;
        loc[5] = cl.z >> 16;
        loc[6] = cl.w & 0xFFFFu;
        loc[7] = cl.w >> 16;
        len[0] = ca.x;
        len[1] = ca.y;
        len[2] = ca.z;
        len[3] = ca.w;
        len[4] = cb.x;
        len[5] = cb.y;
        len[6] = cb.z;
        len[7] = cb.w;
        // Phase 1: the chunk's attenuations (they do not depend on psi),
        // then the next chunk's segment loads are issued.
        float q[kChunk], at[kChunk][NP];
#pragma unroll
        for (int u = 0; u < kChunk; u++) {
            const float2 e = *(const float2 *)((const char *)sq + loc[u]);
            q[u] = e.y;
            const float tau0 = e.x * len[u];
#pragma unroll
            for (int p = 0; p < NP; p++) at[u][p] = attenuation(tau0 * rk[p]);
        }
        // the next chunk's loads, without a branch: the last chunk reloads
        // itself (the pointer stays inside the lane's stream)
        const int adv = c + 1 < nchunk ? nr : 0;
        lp += adv;
        ap += 2 * adv;
        cl = lp[0];
        ca = ap[0];
        cb = ap[1];
        __builtin_amdgcn_sched_barrier(0);
        // Phase 2: the psi recurrence and the tallies. With currents, psi at
        // the start of every step is kept (ph) and the chunk's surface
        // crossings are tallied after it: a wave then runs the tally block
        // once per chunk, not at every step where one of its lanes crosses.
        float ph[kChunk][NP];
#pragma unroll
        for (int u = 0; u < kChunk; u++) {
            if (CUR) {
#pragma unroll
                for (int p = 0; p < NP; p++) ph[u][p] = psi[p];
            }
            float cacc = 0.0f;
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const float d = (psi[p] - q[u]) * at[u][p];
                psi[p] = psi[p] - d;
                cacc = fmaf(d, wt[p], cacc);
            }
            atomicAdd((unsigned long long *)((char *)tal + loc[u]), to_fixed(cacc, sx));
        }
        if (CUR) {
            // crossings at steps t0 .. t0 + 7 use psi at the start of their
            // step (a padding step leaves psi unchanged, so an exit inside
            // the last chunk sees the outgoing flux)
            static_assert(kChunk == 8, "the step select below covers 8 steps");
            const int t0 = c * kChunk;
            while (nstep < t0 + kChunk) {
                const int uu = nstep - t0;
                const bool b0 = uu & 1, b1 = uu & 2, b2 = uu & 4;
                float pe[NP];
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const float a01 = b0 ? ph[1][p] : ph[0][p];
                    const float a23 = b0 ? ph[3][p] : ph[2][p];
                    const float a45 = b0 ? ph[5][p] : ph[4][p];
                    const float a67 = b0 ? ph[7][p] : ph[6][p];
                    const float a03 = b1 ? a23 : a01;
                    const float a47 = b1 ? a67 : a45;
                    pe[p] = b2 ? a47 : a03;
                }
                tally_current(nev, pe);
                next_event();
            }
        }
    }
    if (CUR && active) {
        // events at the exit step (after the last segment)
        while (nstep != 0x10000) {
            tally_current(nev, psi);
            next_event();
        }
    }
    if (active) {
        const int exit_idx = dir ? r.bc0 : r.bc1;
        const int ny = (r.flags >> dir) & 1;
        const int vac = (r.flags >> (2 + dir)) & 1;
        if (!vac) {
#pragma unroll
            for (int p = 0; p < NP; p++) psi_out[(ny ? out_y : out_x) + p * az.stride + exit_idx] = psi[p];
        }
    }
    __syncthreads();
    const double rscale = ldexp(1.0, -kx);
    // each band slot goes to its FSR-major position. All-group sweeps store
    // position-major: the G workgroups of a band run on one XCD, so its L2
    // joins their writes of a slot's groups.
    const int32_t *__restrict__ pos = A.slot_pos + B.loc_base;
    double *__restrict__ pg = A.psum + g * A.psum_sg;
    for (int l = tid; l < B.n_local; l += 2 * W) pg[pos[l] * A.psum_sp] = (double)(long long)tal[l] * rscale;
    if (cur) {
        const double rc = ldexp(1.0, -kc);
        double *__restrict__ cs = A.csum + (size_t)g * A.n_lslots + B.lsurf_base;
        double *__restrict__ ss = A.ssum + (size_t)g * A.n_lslots + B.lsurf_base;
        for (int l = tid; l < B.n_lsurf; l += 2 * W) {
            cs[l] = (double)(long long)tcur[l] * rc;
            ss[l] = (double)(long long)tsfl[l] * rc;
        }
    }
}

// W rays per direction: 2W lanes per workgroup (one wave for W = 32).
template <int NP, bool CUR, int W>
__global__ __launch_bounds__(2 * W) void k_sweep(SweepArgs A)
{
    extern __shared__ unsigned long long smem[];
    __shared__ unsigned int s_bound;
    int g, band;
    if (A.g_fixed >= 0) {
        g = A.g_fixed;
        band = A.order[blockIdx.x];
    } else {
        // XCD-aware order: the G workgroups of a band are consecutive
        // workgroups of one XCD, so they share its L2 copy of the band.
        const int x = blockIdx.x % kNumXcd;
        const int s = blockIdx.x / kNumXcd;
        const int rank = (s / A.G) * kNumXcd + x;
        if (rank >= A.n_bands) return;
        g = s % A.G;
        band = A.order[rank];
    }
    sweep_band<NP, CUR, W>(A, band, g, smem, s_bound);
}

// Persistent work-queue sweep over many problems. Thread 0 takes the next
// item index with a (vector) global atomic and shares it through LDS; every
// workgroup leaves once the counter passes n_items, so the grid always drains.
// The CUR instantiation tallies coarse-surface currents for the problems
// whose want_cur is set (CMFD problems on their last inner iteration). W rays
// per direction: 2W lanes per workgroup, one wave when every problem's bands
// hold at most 32 rays.
template <int NP, bool CUR, int W>
__global__ __launch_bounds__(2 * W) void k_sweep_queue(QueueArgs Q, const SweepArgs *__restrict__ cases)
{
    extern __shared__ unsigned long long smem[];
    __shared__ unsigned int s_bound;
    __shared__ int s_item;
    while (true) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(Q.counter, 1u);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= Q.n_items) break;
        const int4 w = Q.items[it];
        const int c = __builtin_amdgcn_readfirstlane(w.x);
        if (Q.status[2 * c + 1] != 0 || Q.n_inner[c] <= Q.inner) continue;
        SweepArgs A = cases[c];
        if (!CUR || !(Q.want_cur && Q.want_cur[c] && Q.inner == Q.n_inner[c] - 1)) A.lds_lsurf = 0;
        sweep_band<NP, CUR, W>(A, __builtin_amdgcn_readfirstlane(w.y), __builtin_amdgcn_readfirstlane(w.z), smem,
                               s_bound);
        __syncthreads();
    }
}

// phi = T / (St V) + 4 pi q (moc_sweeper_kernel.inc.hpp:165-173). T is the
// sum of the FSR's band partials, stored contiguously ([position][group]):
// kRedLanes lanes per (FSR, group) take the positions s = first + lane,
// first + lane + kRedLanes, ... and an xor butterfly adds the lane sums;
// every lane pairs the same operands, so all lanes hold the same T and the
// order is fixed. The G items of an FSR are adjacent, so a wave's loads
// cover consecutive addresses.
constexpr int kRedLanes = 8;

__device__ __forceinline__ void reduce_flux_item(const ReduceArgs &A, int item, int lane)
{
    int f, g;
    if (A.g_fixed >= 0) {
        f = item;
        g = A.g_fixed;
    } else {
        f = item / A.G;
        g = item % A.G;
    }
    const bool valid = f < A.F;
    const int G = A.G;
    double T = 0.0;
    if (valid) {
        const double *__restrict__ ps = A.psum + g * A.psum_sg;
        const int s1 = A.slot_ptr[f + 1];
        for (int s = A.slot_ptr[f] + lane; s < s1; s += kRedLanes) T += ps[s * A.psum_sp];
    }
#pragma unroll
    for (int o = 1; o < kRedLanes; o <<= 1) T += __shfl_xor(T, o, kRedLanes);
    if (!valid || lane != 0) return;
    const int m = A.fsr_mat[f];
    const double tr = A.mat_tr[m * G + g];
    const double q = (A.src[(size_t)f * G + g] + A.phi[(size_t)f * G + g] * A.mat_ss[m * G + g]) * (1.0 / (tr * FPI));
    A.phi[(size_t)f * G + g] = T / (tr * A.vol[f]) + q * FPI;
}

// Net current and surface flux of a coarse surface: the sum of the band
// slots that tally it, over its area (moc::Current::post_sweep,
// moc_current_worker.hpp:272-318). Surfaces no ray crosses end up zero, as
// after CoarseData::zero_data_radial. Item t covers (surface n0 + t / G,
// group t % G), or surface n0 + t of group g_fixed.
//
// Each surface gets kCurLanes lanes: they take the slots k = first + lane,
// first + lane + kCurLanes, ... and an xor butterfly adds their sums, an
// order fixed by the slot list, so the one-problem and the batched launches
// add in the same order. A 51x51-pin mesh has only ~7900 surfaces per group,
// too few to fill the chip one lane each.
constexpr int kCurLanes = 16;

__device__ __forceinline__ void reduce_current_item(const CurrentArgs &A, int t, int lane)
{
    int s, g;
    if (A.g_fixed >= 0) {
        s = t;
        g = A.g_fixed;
    } else {
        s = t / A.G;
        g = t % A.G;
    }
    const bool valid = s < A.n1 - A.n0;
    s += A.n0;
    double c = 0.0, f = 0.0;
    if (valid)
        for (int k = A.ptr[s] + lane; k < A.ptr[s + 1]; k += kCurLanes) {
            c += A.csum[(size_t)g * A.n_lslots + A.idx[k]];
            f += A.ssum[(size_t)g * A.n_lslots + A.idx[k]];
        }
#pragma unroll
    for (int o = 1; o < kCurLanes; o <<= 1) {
        c += __shfl_xor(c, o, kCurLanes);
        f += __shfl_xor(f, o, kCurLanes);
    }
    if (!valid || lane != 0) return;
    A.cur[(size_t)s * A.G + g] = c / A.area[s];
    A.sfl[(size_t)s * A.G + g] = f / A.area[s];
}

__global__ void k_reduce_current(CurrentArgs A)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    reduce_current_item(A, gid / kCurLanes, gid % kCurLanes);
}

// The same over the problems of a batch whose last inner iteration this is
// and that tally currents (CMFD): block b covers elements [first, first + 256)
// of problem bm[b].c, kCurLanes elements per (surface, group) item. The first
// block of a problem raises its flag that radial currents exist.
__global__ void k_reduce_current_batch(const CurrentArgs *__restrict__ cases, const BlockMap *__restrict__ bm,
                                       const int32_t *__restrict__ status, const int32_t *__restrict__ n_inner,
                                       int inner)
{
    const BlockMap b = bm[blockIdx.x];
    const int c = __builtin_amdgcn_readfirstlane(b.c);
    if (status[2 * c + 1] != 0 || n_inner[c] - 1 != inner) return;
    const CurrentArgs A = cases[c];
    const int first = __builtin_amdgcn_readfirstlane(b.first);
    if (first == 0 && threadIdx.x == 0 && A.radial_flag) *A.radial_flag = 1;
    const int gid = first + threadIdx.x;
    reduce_current_item(A, gid / kCurLanes, gid % kCurLanes);
}

__global__ void k_reduce_flux(ReduceArgs A)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    reduce_flux_item(A, gid / kRedLanes, gid % kRedLanes);
}

// The same over many problems: block b covers elements [first, first + 256)
// of problem bm[b].c (kRedLanes elements per (FSR, group) item). The first
// block of a problem also flips its boundary buffers (A.gmask; 0 for the
// Gauss-Seidel update); the next sweep reads the parity word in a later launch.
__global__ void k_reduce_flux_batch(const ReduceArgs *__restrict__ cases, const BlockMap *__restrict__ bm,
                                    const int32_t *__restrict__ status, const int32_t *__restrict__ n_inner,
                                    int inner)
{
    const BlockMap b = bm[blockIdx.x];
    const int c = __builtin_amdgcn_readfirstlane(b.c);
    if (status[2 * c + 1] != 0 || n_inner[c] <= inner) return;
    const ReduceArgs A = cases[c];
    const int first = __builtin_amdgcn_readfirstlane(b.first);
    if (first == 0 && threadIdx.x == 0 && A.gmask) *A.parity ^= A.gmask;
    const int gid = first + threadIdx.x;
    reduce_flux_item(A, gid / kRedLanes, gid % kRedLanes);
}

__global__ void k_attenuation(const float *__restrict__ tau, float *__restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = attenuation(tau[i] * 1000.0f);
}

}  // namespace

void launch_attenuation(const float *tau, float *out, int n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_attenuation, dim3((n + 255) / 256), dim3(256), 0, s, tau, out, n);
}

size_t sweep_lds_bytes(const SweepArgs &a)
{
    return (size_t)a.lds_local * 16 + (size_t)a.lds_lsurf * 16;
}

void launch_sweep(const SweepArgs &a, int n_polar, bool cur, int n_items, hipStream_t s)
{
    SweepArgs b = a;
    if (!cur) b.lds_lsurf = 0;
    const size_t lds = sweep_lds_bytes(b);
    const bool narrow = a.wg_rays == 32;
    dim3 grid(n_items), block(narrow ? 64 : kWG);
#define MOCGPU_SWEEP_CASE(NPV)                                                                 \
    case NPV:                                                                                  \
        if (narrow) {                                                                          \
            if (cur) hipLaunchKernelGGL((k_sweep<NPV, true, 32>), grid, block, lds, s, b);     \
            else hipLaunchKernelGGL((k_sweep<NPV, false, 32>), grid, block, lds, s, b);        \
        } else {                                                                               \
            if (cur) hipLaunchKernelGGL((k_sweep<NPV, true, kRays>), grid, block, lds, s, b);  \
            else hipLaunchKernelGGL((k_sweep<NPV, false, kRays>), grid, block, lds, s, b);     \
        }                                                                                      \
        break;
    switch (n_polar) {
        MOCGPU_SWEEP_CASE(1)
        MOCGPU_SWEEP_CASE(2)
        MOCGPU_SWEEP_CASE(3)
        MOCGPU_SWEEP_CASE(4)
    }
#undef MOCGPU_SWEEP_CASE
}

int queue_blocks_per_cu(int n_polar, size_t lds, bool cur, int wg_rays)
{
    int nb = 0;
    hipError_t e = hipSuccess;
    const bool narrow = wg_rays == 32;
#define MOCGPU_QOCC(NPV)                                                                                         \
    case NPV:                                                                                                    \
        if (narrow)                                                                                              \
            e = cur ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep_queue<NPV, true, 32>, 64, lds)   \
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep_queue<NPV, false, 32>, 64, lds); \
        else                                                                                                     \
            e = cur ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep_queue<NPV, true, kRays>, kWG, lds) \
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep_queue<NPV, false, kRays>, kWG,   \
                                                                   lds);                                         \
        break;
    switch (n_polar) {
        MOCGPU_QOCC(1)
        MOCGPU_QOCC(2)
        MOCGPU_QOCC(3)
    default:
        MOCGPU_QOCC(4)
    }
#undef MOCGPU_QOCC
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

void launch_sweep_queue(const QueueArgs &q, int n_polar, size_t lds, int n_blocks, hipStream_t s, bool cur,
                        int wg_rays)
{
    const bool narrow = wg_rays == 32;
    dim3 grid(n_blocks), block(narrow ? 64 : kWG);
#define MOCGPU_QLAUNCH(NPV)                                                                                  \
    case NPV:                                                                                                \
        if (narrow) {                                                                                        \
            if (cur) hipLaunchKernelGGL((k_sweep_queue<NPV, true, 32>), grid, block, lds, s, q, q.cases);    \
            else hipLaunchKernelGGL((k_sweep_queue<NPV, false, 32>), grid, block, lds, s, q, q.cases);       \
        } else {                                                                                             \
            if (cur) hipLaunchKernelGGL((k_sweep_queue<NPV, true, kRays>), grid, block, lds, s, q, q.cases); \
            else hipLaunchKernelGGL((k_sweep_queue<NPV, false, kRays>), grid, block, lds, s, q, q.cases);    \
        }                                                                                                    \
        break;
    switch (n_polar) {
        MOCGPU_QLAUNCH(1)
        MOCGPU_QLAUNCH(2)
        MOCGPU_QLAUNCH(3)
    default:
        MOCGPU_QLAUNCH(4)
    }
#undef MOCGPU_QLAUNCH
}

int sweep_grid_items(int n_bands, int G, bool all_groups)
{
    if (!all_groups) return n_bands;
    const int per_xcd = (n_bands + kNumXcd - 1) / kNumXcd;
    return per_xcd * G * kNumXcd;
}

void launch_reduce_current(const CurrentArgs &a, size_t n_items, hipStream_t s)
{
    if (n_items > 0)
        hipLaunchKernelGGL(k_reduce_current, dim3((n_items * kCurLanes + 255) / 256), dim3(256), 0, s, a);
}

void launch_reduce_current_batch(const CurrentArgs *cases, const BlockMap *bm, int n_blocks, const int32_t *status,
                                 const int32_t *n_inner, int inner, hipStream_t s)
{
    if (n_blocks > 0)
        hipLaunchKernelGGL(k_reduce_current_batch, dim3(n_blocks), dim3(256), 0, s, cases, bm, status, n_inner, inner);
}

void launch_reduce_flux_batch(const ReduceArgs *cases, const BlockMap *bm, int n_blocks, const int32_t *status,
                              const int32_t *n_inner, int inner, hipStream_t s)
{
    if (n_blocks > 0)
        hipLaunchKernelGGL(k_reduce_flux_batch, dim3(n_blocks), dim3(256), 0, s, cases, bm, status, n_inner, inner);
}

int reduce_flux_lanes() { return kRedLanes; }
int reduce_current_lanes() { return kCurLanes; }

void launch_reduce_flux(const ReduceArgs &a, hipStream_t s)
{
    const size_t n = (size_t)(a.g_fixed >= 0 ? a.F : a.F * a.G) * kRedLanes;
    hipLaunchKernelGGL(k_reduce_flux, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
}

}  // namespace gpu
}  // namespace mocc_amd
