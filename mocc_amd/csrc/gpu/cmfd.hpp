// Pin-cell CMFD on the device (FP64): homogenization, D-tilde/D-hat coupling,
// the coarse power iteration and its BiCGSTAB inner solves.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../../include/mocgpu.h"

namespace mocc_amd {
namespace gpu {

// Device arrays of the sweeper the CMFD reads and writes (owned by the
// sweeper handle). Layouts: phi [F][G], cur / sfl [n_surf][G].
struct CmfdView {
    int F, G, npin, n_mat;
    int nc;  // coarse cells of the plane (nx * ny)
    const int32_t *fsr_mat, *pin_off, *pin_cell;
    const double *area, *vol, *pin_area;
    const double *mat_tr, *mat_nf, *mat_f, *mat_ch, *mat_sc;  // [m][g], sc [m][to][from]
    double *phi, *cur, *sfl;
    double *k;  // device scalar, in/out
    int has_radial;
    int neg_fixup;  // clamp negative coarse fluxes to 0 before the solve (<cmfd negative_fixup>)
};

class CmfdSolver {
public:
    // Largest coarse mesh of a single problem's solve (1024 lanes x 6 cells),
    // and of a problem in a batch (512 lanes x 6 cells).
    static constexpr int kMaxCells = 6144;
    static constexpr int kMaxBatchCells = 3072;

    ~CmfdSolver();
    void setup(const mocgpu_problem &p, const mocgpu_options &opt, hipStream_t s);
    void initialize(hipStream_t s);
    // The groups that scatter into each group (from the material cross
    // sections); called by setup and again when the cross sections change.
    void update_xs(const mocgpu_problem &p, hipStream_t s);
    // EigenSolver::do_cmfd (eigen_solver.cpp:263-291) enqueued on s: coarse
    // flux from the fine flux, CMFD::solve(k) with k on the device, then the
    // fine flux rescaled pin by pin. Nothing is read back.
    void enqueue_step(const CmfdView &v, hipStream_t s);
    // Outer iterations of the last solve and the non-positive-flux flag of
    // the last rescale (read back; synchronizes s).
    // iterations of the last solve: coarse power iterations, the rescale's
    // non-positive flux flag, BiCGSTAB iterations summed over all groups
    void read_info(hipStream_t s, int *iterations, int *bad_flux, int *bicgstab_iterations = nullptr);

    int n_cell() const { return nc_; }
    // device word raised when the last rescale met a non-positive coarse flux
    const int *bad_flux_flag() const { return d_info_ ? d_info_ + 1 : nullptr; }
    // Assembled coefficients of group g after the last solve, rows in cell
    // order with the columns of each row sorted (CSR, for tests).
    void matrix_csr(int g, hipStream_t s, std::vector<int32_t> &row_ptr, std::vector<int32_t> &col,
                    std::vector<double> &val);
    void coarse_flux(hipStream_t s, std::vector<double> &out);
    // XSMeshHomogenized::update (xs_mesh_homogenized.cpp:176-197) alone: the
    // pin-cell cross sections from the current fine flux, as the 2-D/3-D
    // sweeper updates its Sn cross-section mesh before its correction sweep.
    void homogenize(const CmfdView &v, hipStream_t s);
    // homogenized transport cross section [n_cell][G] of the last homogenization
    const double *homog_tr() const { return d_htr_; }
    // CoarseData::partial_current / partial_current_old [n_surf][G][2] of the
    // latest and the previous solve (cmfd.cpp:583-599), and whether the old
    // ones exist (set_has_old_partial(n_solve > 0) at the store)
    const double *partial() const { return d_part_; }
    const double *partial_old() const { return d_part_old_; }
    bool has_old_partial() const { return has_old_partial_; }

private:
    friend struct CmfdAccess;
    void free_all();
    int nx_ = 0, ny_ = 0, nc_ = 0, G_ = 0, n_surf_ = 0;
    double hz_ = 0.0;
    int bc_[6] = {0, 0, 0, 0, 0, 0};
    mocgpu_options opt_{};
    double *d_dx_ = nullptr, *d_dy_ = nullptr, *d_cvol_ = nullptr;
    double *d_cflux_ = nullptr;                                             // [G][nc]
    uint64_t *d_smask_ = nullptr;  // [G]: bit g' set when some material scatters from g' into g (g' != g)
    double *d_htr_ = nullptr, *d_hnf_ = nullptr, *d_hf_ = nullptr, *d_hch_ = nullptr, *d_hrm_ = nullptr;  // [nc][G]
    double *d_hsc_ = nullptr;                                               // [nc][G][G]
    double *d_dtil_ = nullptr, *d_stil_ = nullptr, *d_dhat_ = nullptr, *d_shat_ = nullptr;  // [n_surf][G]
    double *d_coef_ = nullptr;  // [5][G][nc]: south, west, diagonal, east, north
    double *d_gscr_ = nullptr;  // [3][nc]: previous coarse fission source, BiCGSTAB right-hand side and inverse diagonal (solves with the rows in LDS)
                                // (solves that keep the rows in LDS)
    int *d_info_ = nullptr;     // [0] outer iterations, [1] non-positive pin flux seen
    double *d_part_ = nullptr, *d_part_old_ = nullptr;  // [n_surf][G][2] partial currents
    int n_solve_ = 0;
    bool has_old_partial_ = false;
};

// One problem of a batch that runs CMFD: its index in the batch's status
// array, its solver, the view of its sweeper's arrays (k and the fine flux
// on the device) and a device flag that turns 1 once radial currents exist.
struct CmfdBatchEntry {
    int batch_index;
    CmfdSolver *solver;
    CmfdView view;
    int32_t *radial_flag;
};

// The CMFD step (CmfdSolver::enqueue_step) of many problems in a fixed
// sequence of launches, each covering every problem: the element-wise
// set-up kernels over (problem, element) block maps and one solve
// workgroup per problem. Problems whose status is finished are skipped, so
// the sequence can be captured once and replayed.
class CmfdBatch {
public:
    CmfdBatch();
    ~CmfdBatch();
    CmfdBatch(const CmfdBatch &) = delete;
    CmfdBatch &operator=(const CmfdBatch &) = delete;
    void build(const std::vector<CmfdBatchEntry> &entries, hipStream_t s);
    void enqueue(const int32_t *status, hipStream_t s) const;
    bool empty() const { return n_ == 0; }

private:
    struct Impl;
    Impl *impl_ = nullptr;
    int n_ = 0;
};

}  // namespace gpu
}  // namespace mocc_amd
