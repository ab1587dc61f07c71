// Batch engine: many independent k-eigenvalue problems (cross-section branch
// cases, lattices of mixed size) advanced together on one device.
//
// One outer iteration of EigenSolver::step (eigen_solver.cpp:225-253) for
// every unfinished problem is a fixed sequence of launches, each covering all
// problems at once:
//
//   k_b_pre: fission source, phi_old and the group sources (Jacobi in energy)
//   n_inner x [ work-queue sweep of every (problem, band, group) item (two
//               queues for Gauss-Seidel problems: octant-1 bands, then
//               octant-2 bands) -> flux reduction + boundary parity flip ]
//   k_b_post: total fission and k, normalization, err_psi and the
//             convergence test, one workgroup per problem.
//
// Problems with CMFD first run the CMFD step (CmfdBatch: coarse flux,
// homogenization, coupling, assembly, one solve workgroup per problem,
// currents and the pin rescale), and their last inner sweep tallies
// coarse-surface currents, reduced per problem before k_b_post.
//
// The sequence is captured once per batch composition into a HIP graph and
// replayed; problems that have converged stay in the graph and are skipped by
// every kernel from their status word. Each problem's arithmetic (fixed-point
// band tallies, fixed-order flux sums, one workgroup per problem for every
// reduction) does not depend on the other problems, so a problem's result is
// bitwise the same alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "handle.hpp"

namespace {

struct CaseDesc {
    int32_t F, G, max_iter, min_iter;
    double k_tol, psi_tol;
    const int *cmfd_bad;  // the CMFD rescale's non-positive flux flag, or null
    const int32_t *fsr_mat;
    const double *mat_nf, *mat_ch, *mat_sc, *vol;
    double *phi, *phi_old, *src, *fs, *fs_prev;
    double *scal;         // [S_COUNT]
};

constexpr int kBT = 256;  // threads per block of the element-wise launches

__device__ __forceinline__ bool finished(const int32_t *status, int c) { return status[2 * c + 1] != 0; }

// The start of an outer iteration for one FSR of every unfinished problem:
// fission source fs = (1/k) sum_g nuSf phi and fs_prev = fs
// (transport_sweeper.cpp:119-134), phi_old = phi, and every group's source
// S = chi fs + sum_{g' != g} Ss[g <- g'] phi_old (source.cpp:41-112).
// Block 0 also zeroes the work-queue counters of the iteration's sweeps.
__global__ void k_b_pre(const CaseDesc *__restrict__ cs, const BlockMap *__restrict__ bm,
                        const int32_t *__restrict__ status, unsigned int *__restrict__ counters, int n_counters)
{
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < n_counters; i += blockDim.x) counters[i] = 0u;
    const BlockMap b = bm[blockIdx.x];
    const int c = __builtin_amdgcn_readfirstlane(b.c);
    if (finished(status, c)) return;
    const CaseDesc &d = cs[c];
    const int f = b.first + threadIdx.x;
    if (f >= d.F) return;
    const int G = d.G;
    const int m = d.fsr_mat[f];
    const double rk = 1.0 / d.scal[S_K];
    const double *phi = d.phi + (size_t)f * G;
    double *phi_old = d.phi_old + (size_t)f * G;
    double fs = 0.0;
    for (int g = 0; g < G; g++) {
        const double v = phi[g];
        fs += rk * d.mat_nf[m * G + g] * v;
        phi_old[g] = v;
    }
    d.fs[f] = fs;
    d.fs_prev[f] = fs;
    for (int g = 0; g < G; g++) {
        double s = d.mat_ch[m * G + g] * fs;
        const double *row = d.mat_sc + ((size_t)m * G + g) * G;
        for (int gg = 0; gg < G; gg++) {
            if (gg == g) continue;
            s += row[gg] * phi_old[gg];
        }
        d.src[(size_t)f * G + g] = s;
    }
}

constexpr int kPostT = 1024;  // threads of k_b_post: one workgroup per problem
constexpr int kPostWaves = kPostT / 64;

// Sum of n values over the workgroup in a fixed order: an xor butterfly in
// each wave, then every lane adds the wave sums in wave order. red holds
// kPostWaves x 4 doubles; the leading barrier lets a caller reuse it.
__device__ __forceinline__ void block_sum_n(double *acc, int n, double (*red)[4])
{
    for (int q = 0; q < n; q++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < n; q++) red[threadIdx.x >> 6][q] = acc[q];
    __syncthreads();
    for (int q = 0; q < n; q++) {
        double s = 0.0;
        for (int w = 0; w < kPostWaves; w++) s += red[w][q];
        acc[q] = s;
    }
}

// The end of an outer iteration, one workgroup per problem (fixed reduction
// trees): total fission and k <- k F_new / F_old (eigen_solver.cpp:242-247),
// the new fission source, Normalize of fs and fs_prev (util/utils.hpp:28-51),
// err_psi (eigen_solver.cpp:171-184) and EigenSolver::solve's test:
// status[2c] counts steps, status[2c+1] becomes 1 (converged), 2 (max_iter)
// or 3 (k not a number).
__global__ __launch_bounds__(kPostT) void k_b_post(const CaseDesc *__restrict__ cs, int32_t *__restrict__ status)
{
    const int c = blockIdx.x;
    if (finished(status, c)) return;
    const CaseDesc &d = cs[c];
    __shared__ double sh[kPostWaves][4];
    const int G = d.G;
    double *s = d.scal;
    double acc[4] = {0, 0, 0, 0};
    for (int f = threadIdx.x; f < d.F; f += kPostT) {
        const int m = d.fsr_mat[f];
        for (int g = 0; g < G; g++) {
            acc[0] += d.phi[(size_t)f * G + g] * d.vol[f] * d.mat_nf[m * G + g];
            acc[1] += d.phi_old[(size_t)f * G + g] * d.vol[f] * d.mat_nf[m * G + g];
        }
    }
    block_sum_n(acc, 2, sh);
    // keff_prev_ is k after the step's CMFD solve (eigen_solver.cpp:246-247)
    const double k_prev = s[S_K];
    const double k = k_prev * acc[0] / acc[1];
    const double rk = 1.0 / k;
    double nrm[4] = {0, 0, 0, 0};
    for (int f = threadIdx.x; f < d.F; f += kPostT) {
        const int m = d.fsr_mat[f];
        double v = 0.0;
        for (int g = 0; g < G; g++) v += rk * d.mat_nf[m * G + g] * d.phi[(size_t)f * G + g];
        d.fs[f] = v;
        const double vp = d.fs_prev[f];
        nrm[0] += v > 0.0 ? 1.0 : 0.0;
        nrm[1] += v;
        nrm[2] += vp > 0.0 ? 1.0 : 0.0;
        nrm[3] += vp;
    }
    block_sum_n(nrm, 4, sh);
    const double fa = nrm[0] / nrm[1], fb = nrm[2] / nrm[3];
    double e2[1] = {0.0};
    for (int f = threadIdx.x; f < d.F; f += kPostT) {
        const double e = d.fs[f] * fa - d.fs_prev[f] * fb;
        e2[0] += e * e;
    }
    block_sum_n(e2, 1, sh);
    if (threadIdx.x != 0) return;
    s[S_TF_NEW] = acc[0];
    s[S_TF_OLD] = acc[1];
    s[S_KPREV] = k_prev;
    s[S_K] = k;
    s[S_ERRK] = fabs(k - k_prev);
    s[S_FA] = fa;
    s[S_FB] = fb;
    s[S_ERRPSI] = sqrt(e2[0] / s[S_NFISS]);
    const int it = ++status[2 * c];
    int fin = 0;
    if (d.cmfd_bad && *d.cmfd_bad) fin = 4;  // set_pin_flux_1g throws (moc_sweeper.cpp:399-402)
    else if (!(k == k)) fin = 3;
    else if (s[S_ERRK] < d.k_tol && s[S_ERRPSI] < d.psi_tol && it - 1 >= d.min_iter) fin = 1;
    else if (it >= d.max_iter) fin = 2;
    status[2 * c + 1] = fin;
}

std::vector<BlockMap> block_map(const std::vector<size_t> &counts)
{
    std::vector<BlockMap> m;
    for (size_t c = 0; c < counts.size(); c++)
        for (size_t first = 0; first < counts[c]; first += kBT) m.push_back(BlockMap{(int32_t)c, (int32_t)first});
    return m;
}

}  // namespace

struct mocgpu_batch {
    std::string err;
    hipStream_t stream = nullptr;
    int64_t budget = 0, used = 0;
    bool use_graph = true;
    struct Case {
        mocgpu_handle *h;
        int32_t id;
        double k_tol, psi_tol;
        int32_t max_iter, min_iter;
        int32_t status[2];  // host mirror, valid after a poll
        bool reported;
        int64_t bytes;
    };
    std::vector<Case> cases;
    int32_t next_id = 1;
    bool dirty = true;
    int NP = 0;
    // device tables of the current composition
    DBuf<CaseDesc> d_desc;
    DBuf<SweepArgs> d_sweep;
    DBuf<ReduceArgs> d_red;
    DBuf<int4> d_items;
    DBuf<BlockMap> d_bmF, d_bmR, d_bmJ;
    DBuf<int32_t> d_status, d_ninner, d_wantcur;
    DBuf<CurrentArgs> d_cura;
    DBuf<uint32_t> d_counter;
    int n_bmF = 0, n_bmR = 0, n_bmJ = 0, n_items = 0, max_inner = 1, q_blocks = 1, q_blocks_cur = 1;
    int wg_rays = kRays;  // rays per direction of a queue workgroup (32 or 64)
    int phase_first[2] = {0, 0}, phase_items[2] = {0, 0};
    size_t lds = 0, lds_cur = 0;
    // CMFD problems: their CMFD step runs first in every outer iteration, and
    // the inner iterations marked in cur_inner tally currents
    CmfdBatch cmfd;
    std::vector<char> cur_inner;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~mocgpu_batch()
    {
        // the members are the caller's again once the batch's work is done
        if (stream) (void)hipStreamSynchronize(stream);
        for (const Case &c : cases) c.h->batch = nullptr;
        drop_graph();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    void sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

    // status of every case from the device into the host mirrors
    void pull_status()
    {
        if (!d_status.p || cases.empty()) return;
        std::vector<int32_t> st(2 * cases.size());
        HIP_CHECK(hipMemcpyAsync(st.data(), d_status.p, st.size() * 4, hipMemcpyDeviceToHost, stream));
        sync();
        for (size_t i = 0; i < cases.size(); i++) {
            cases[i].status[0] = st[2 * i];
            cases[i].status[1] = st[2 * i + 1];
        }
    }

    void rebuild()
    {
        drop_graph();
        const int n = (int)cases.size();
        std::vector<CaseDesc> desc(n);
        std::vector<SweepArgs> sw(n);
        std::vector<ReduceArgs> rd(n);
        std::vector<int32_t> st(2 * n), ni(n);
        std::vector<size_t> cF(n), cR(n), cJ(n, 0);
        std::vector<int32_t> wc(n, 0);
        std::vector<CurrentArgs> ja(n);
        std::vector<CmfdBatchEntry> cm;
        struct It {
            int64_t cost;
            int4 w;
        };
        std::vector<It> items[2];
        lds = 0;
        lds_cur = 0;
        max_inner = 1;
        for (int i = 0; i < n; i++) {
            mocgpu_handle *h = cases[i].h;
            CaseDesc &d = desc[i];
            d.F = h->F;
            d.G = h->G;
            d.max_iter = cases[i].max_iter;
            d.min_iter = cases[i].min_iter;
            d.k_tol = cases[i].k_tol;
            d.psi_tol = cases[i].psi_tol;
            d.fsr_mat = h->d_fsr_mat.p;
            d.mat_nf = h->d_mat_nf.p;
            d.mat_ch = h->d_mat_ch.p;
            d.mat_sc = h->d_mat_sc.p;
            d.vol = h->d_vol.p;
            d.phi = h->d_phi.p;
            d.phi_old = h->d_phi_old.p;
            d.src = h->d_src.p;
            d.fs = h->d_fs.p;
            d.fs_prev = h->d_fs_prev.p;
            d.scal = h->d_scal.p;
            d.cmfd_bad = nullptr;
            SweepArgs a = h->sweep_args(-1);
            a.parity_dev = h->d_parity.p;
            if (h->opt.use_cmfd) {
                // CMFD: currents on the last inner iteration, the CMFD step first
                wc[i] = 1;
                d.cmfd_bad = h->cmfd->bad_flux_flag();
                lds_cur = std::max(lds_cur, sweep_lds_bytes(a));
                ja[i] = h->current_args(-1);
                ja[i].radial_flag = h->d_radial.p;
                cJ[i] = (size_t)(ja[i].n1 - ja[i].n0) * h->G * reduce_current_lanes();
                cm.push_back(CmfdBatchEntry{i, h->cmfd.get(), h->cmfd_view(), h->d_radial.p});
            } else {
                a.lds_lsurf = 0;
            }
            sw[i] = a;
            SweepArgs plain = a;
            plain.lds_lsurf = 0;
            lds = std::max(lds, sweep_lds_bytes(plain));
            lds_cur = std::max(lds_cur, sweep_lds_bytes(plain));
            rd[i] = h->reduce_args(-1);
            // Gauss-Seidel problems keep one boundary buffer: nothing to flip
            rd[i].parity = h->d_parity.p;
            rd[i].gmask = h->opt.gs_boundary ? 0ull : (h->G >= 64 ? ~0ull : ((1ull << h->G) - 1ull));
            st[2 * i] = cases[i].status[0];
            st[2 * i + 1] = cases[i].status[1];
            ni[i] = std::max(1, h->opt.n_inner);
            max_inner = std::max(max_inner, ni[i]);
            cF[i] = (size_t)h->F;
            cR[i] = (size_t)h->F * h->G * reduce_flux_lanes();
            // phase 0: every band (Jacobi) or the octant-1 bands (Gauss-Seidel);
            // phase 1: the octant-2 bands of Gauss-Seidel problems
            const SweepPlan &pl = *h->plan;
            for (size_t r = 0; r < pl.band_order.size(); r++) {
                const int band = pl.band_order[r];
                const int64_t cost = (int64_t)pl.bands[band].nmax;
                const int ph = (h->opt.gs_boundary && pl.bands[band].azi >= pl.A / 2) ? 1 : 0;
                for (int g = 0; g < h->G; g++) items[ph].push_back(It{cost, make_int4(i, band, g, 0)});
            }
        }
        // per phase, longest items first over all problems (the G items of a
        // band stay together)
        std::vector<int4> iv;
        for (int ph = 0; ph < 2; ph++) {
            std::stable_sort(items[ph].begin(), items[ph].end(), [](const It &x, const It &y) { return x.cost > y.cost; });
            phase_first[ph] = (int)iv.size();
            for (const It &it : items[ph]) iv.push_back(it.w);
            phase_items[ph] = (int)items[ph].size();
        }
        n_items = (int)iv.size();
        auto bF = block_map(cF), bR = block_map(cR), bJ = block_map(cJ);
        n_bmF = (int)bF.size();
        n_bmR = (int)bR.size();
        n_bmJ = (int)bJ.size();
        // inner iterations on which some CMFD problem tallies currents (its last)
        cur_inner.assign(max_inner, 0);
        for (int i = 0; i < n; i++)
            if (wc[i]) cur_inner[ni[i] - 1] = 1;
        d_desc.upload(desc, stream);
        d_sweep.upload(sw, stream);
        d_red.upload(rd, stream);
        d_items.upload(iv, stream);
        d_bmF.upload(bF, stream);
        d_bmR.upload(bR, stream);
        d_status.upload(st, stream);
        d_ninner.upload(ni, stream);
        d_wantcur.upload(wc, stream);
        d_cura.upload(ja, stream);
        d_bmJ.upload(bJ, stream);
        cmfd.build(cm, stream);
        d_counter.alloc(2 * (size_t)max_inner);  // one work-queue counter per (inner, phase)
        int dev = 0, ncu = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        // one-wave workgroups when every problem's (possibly autotuned) band
        // plan holds at most 32 rays per band; a band's arithmetic is the
        // same in either width
        wg_rays = 32;
        for (int i = 0; i < n; i++)
            if (cases[i].h->plan->max_band_rays > 32) wg_rays = kRays;
        q_blocks = std::max(1, ncu * queue_blocks_per_cu(NP, lds, false, wg_rays));
        q_blocks_cur = cm.empty() ? q_blocks : std::max(1, ncu * queue_blocks_per_cu(NP, lds_cur, true, wg_rays));
        sync();
        dirty = false;
    }

    // one outer iteration of every unfinished problem
    void enqueue_outer()
    {
        const int n = (int)cases.size();
        hipStream_t s = stream;
        // EigenSolver::step: do_cmfd first (eigen_solver.cpp:228-230)
        cmfd.enqueue(d_status.p, s);
        hipLaunchKernelGGL(k_b_pre, dim3(n_bmF), dim3(kBT), 0, s, d_desc.p, d_bmF.p, d_status.p, d_counter.p,
                           2 * max_inner);
        for (int inner = 0; inner < max_inner; inner++) {
            const bool cur = cur_inner[inner] != 0;
            for (int ph = 0; ph < 2; ph++) {
                if (phase_items[ph] == 0) continue;
                QueueArgs q{};
                q.cases = d_sweep.p;
                q.items = d_items.p + phase_first[ph];
                q.n_items = phase_items[ph];
                q.inner = inner;
                q.counter = d_counter.p + 2 * inner + ph;
                q.status = d_status.p;
                q.n_inner = d_ninner.p;
                q.want_cur = d_wantcur.p;
                launch_sweep_queue(q, NP, cur ? lds_cur : lds, std::min(cur ? q_blocks_cur : q_blocks, phase_items[ph]),
                                   s, cur, wg_rays);
            }
            launch_reduce_flux_batch(d_red.p, d_bmR.p, n_bmR, d_status.p, d_ninner.p, inner, s);
            if (cur) launch_reduce_current_batch(d_cura.p, d_bmJ.p, n_bmJ, d_status.p, d_ninner.p, inner, s);
        }
        hipLaunchKernelGGL(k_b_post, dim3(n), dim3(kPostT), 0, s, d_desc.p, d_status.p);
    }

    void run(int n_outer)
    {
        if (cases.empty() || n_outer <= 0) return;
        if (dirty) rebuild();
        if (use_graph) {
            if (!exec) {
                HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                enqueue_outer();
                HIP_CHECK(hipStreamEndCapture(stream, &graph));
                HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            }
            for (int i = 0; i < n_outer; i++) HIP_CHECK(hipGraphLaunch(exec, stream));
        } else {
            for (int i = 0; i < n_outer; i++) enqueue_outer();
            HIP_CHECK(hipGetLastError());
        }
    }

    int find(int32_t id) const
    {
        for (size_t i = 0; i < cases.size(); i++)
            if (cases[i].id == id) return (int)i;
        return -1;
    }
};

namespace {
template <typename F> int bguard(mocgpu_batch *b, F &&f)
{
    try {
        return f();
    } catch (const HipError &e) {
        if (b) b->err = e.what();
        return MOCGPU_ERR_DEVICE;
    } catch (const std::exception &e) {
        if (b) b->err = e.what();
        return MOCGPU_ERR_ARG;
    }
}
thread_local std::string g_batch_err;
}  // namespace

extern "C" {

int mocgpu_batch_create(int64_t budget_bytes, int32_t use_graph, mocgpu_batch **out)
{
    if (!out) return MOCGPU_ERR_ARG;
    auto b = std::make_unique<mocgpu_batch>();
    int rc = bguard(b.get(), [&] {
        HIP_CHECK(hipStreamCreate(&b->stream));
        b->budget = budget_bytes;
        b->use_graph = use_graph != 0;
        return 0;
    });
    if (rc) {
        g_batch_err = b->err;
        return rc;
    }
    *out = b.release();
    return 0;
}

int mocgpu_batch_destroy(mocgpu_batch *b)
{
    delete b;
    return 0;
}

const char *mocgpu_batch_last_error(const mocgpu_batch *b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int mocgpu_batch_add(mocgpu_batch *b, mocgpu_handle *h, double k_tol, double psi_tol, int32_t max_iter,
                     int32_t min_iter, int32_t *id)
{
    if (!b || !h || !id) return MOCGPU_ERR_ARG;
    return bguard(b, [&] {
        // The batch sweeps every group of a problem in one launch from the
        // previous outer iteration's flux: Jacobi in energy only. A handle in
        // mocc's Gauss-Seidel energy mode (energy_mode 0) would follow another
        // iteration path than its own mocgpu_eigen_solve.
        if (h->opt.energy_mode != 1) {
            b->err = "batched problems sweep Jacobi in energy: create the handle with energy_mode = 1";
            return (int)MOCGPU_ERR_UNSUPPORTED;
        }
        if (h->opt.use_cmfd && h->cmfd && h->cmfd->n_cell() > CmfdSolver::kMaxBatchCells) {
            b->err = "a batched CMFD problem holds at most " + std::to_string(CmfdSolver::kMaxBatchCells) +
                     " coarse cells (this one has " + std::to_string(h->cmfd->n_cell()) + ")";
            return (int)MOCGPU_ERR_UNSUPPORTED;
        }
        if (!b->cases.empty() && h->NP != b->NP) {
            b->err = "all problems of a batch must have the same number of polar angles";
            return (int)MOCGPU_ERR_UNSUPPORTED;
        }
        if (h->batch == b) throw std::runtime_error("handle already in the batch");
        if (h->batch) throw std::runtime_error("handle is a member of another batch");
        // the host status mirrors must be current before the composition changes
        b->sync();
        if (!b->dirty) b->pull_status();
        const int64_t bytes = mocgpu_state_bytes(h);
        if (b->budget > 0 && b->used + bytes > b->budget) {
            b->err = "device memory budget exhausted";
            return (int)MOCGPU_ERR_CAPACITY;
        }
        // EigenSolver::solve's start (mocgpu_eigen_solve): flat flux, the
        // initial incoming flux, k = 1, step counter 0
        int rc = mocgpu_initialize(h);
        if (rc) {
            b->err = h->err;
            return rc;
        }
        std::vector<double> sc(S_COUNT, 0.0);
        sc[S_K] = 1.0;
        sc[S_KPREV] = 1.0;
        sc[S_NFISS] = (double)h->n_fissile;
        HIP_CHECK(hipMemcpyAsync(h->d_scal.p, sc.data(), sc.size() * 8, hipMemcpyHostToDevice, b->stream));
        const uint64_t zero = 0;
        HIP_CHECK(hipMemcpyAsync(h->d_parity.p, &zero, 8, hipMemcpyHostToDevice, b->stream));
        const int32_t no_radial = 0;  // the first CMFD solve has no MoC currents yet (D-hat = 0)
        HIP_CHECK(hipMemcpyAsync(h->d_radial.p, &no_radial, 4, hipMemcpyHostToDevice, b->stream));
        b->sync();
        mocgpu_batch::Case c{};
        c.h = h;
        c.id = b->next_id++;
        c.k_tol = k_tol;
        c.psi_tol = psi_tol;
        c.max_iter = max_iter;
        c.min_iter = min_iter;
        c.bytes = bytes;
        b->cases.push_back(c);
        b->used += bytes;
        b->NP = h->NP;
        b->dirty = true;
        h->batch = b;
        *id = c.id;
        return 0;
    });
}

int mocgpu_batch_run(mocgpu_batch *b, int32_t n_outer)
{
    if (!b) return MOCGPU_ERR_ARG;
    return bguard(b, [&] {
        b->run(n_outer);
        return 0;
    });
}

int mocgpu_batch_poll(mocgpu_batch *b, int32_t *ids, mocgpu_eigen_state *states, int32_t cap)
{
    if (!b) return MOCGPU_ERR_ARG;
    return bguard(b, [&] {
        b->sync();
        if (!b->dirty) b->pull_status();
        int n = 0;
        for (auto &c : b->cases) {
            if (c.reported || c.status[1] == 0 || n >= cap) continue;
            double sc[S_COUNT];
            HIP_CHECK(hipMemcpyAsync(sc, c.h->d_scal.p, sizeof(sc), hipMemcpyDeviceToHost, b->stream));
            b->sync();
            if (ids) ids[n] = c.id;
            if (states) {
                mocgpu_eigen_state s{};
                s.k = sc[S_K];
                s.k_prev = sc[S_KPREV];
                s.err_k = sc[S_ERRK];
                s.err_psi = sc[S_ERRPSI];
                s.iterations = c.status[0];
                s.converged = c.status[1] == 1 ? 1 : 0;
                states[n] = s;
            }
            c.reported = true;
            n++;
        }
        return n;
    });
}

int mocgpu_batch_remove(mocgpu_batch *b, int32_t id)
{
    if (!b) return MOCGPU_ERR_ARG;
    return bguard(b, [&] {
        const int i = b->find(id);
        if (i < 0) throw std::runtime_error("no such problem in the batch");
        b->sync();
        if (!b->dirty) b->pull_status();
        b->used -= b->cases[i].bytes;
        b->cases[i].h->batch = nullptr;
        b->cases.erase(b->cases.begin() + i);
        b->dirty = true;
        return 0;
    });
}

int32_t mocgpu_batch_size(const mocgpu_batch *b) { return b ? (int32_t)b->cases.size() : 0; }

int64_t mocgpu_batch_bytes_used(const mocgpu_batch *b) { return b ? b->used : 0; }

}  // extern "C"

void batch_sync(mocgpu_batch *b) { b->sync(); }

void batch_detach(mocgpu_batch *b, mocgpu_handle *h)
{
    for (const auto &c : b->cases)
        if (c.h == h) {
            (void)mocgpu_batch_remove(b, c.id);
            break;
        }
    h->batch = nullptr;
}
