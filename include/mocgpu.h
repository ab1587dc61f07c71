/*
 * mocgpu.h -- C ABI of the MI355X (gfx950) 2-D Method-of-Characteristics
 * sweeper, source update and pin-cell CMFD solve.
 *
 * This is the drop-in boundary for mocc's transport-sweeper plugin
 * interface. Every entry point below replaces one member of the C++ class
 * mocc::TransportSweeper / mocc::moc::MoCSweeper / mocc::CMFD /
 * mocc::EigenSolver (file:line cited per function). The signatures use only
 * plain C types and pointers so that any host language can bind them; a
 * handle is not thread-safe, distinct handles are independent.
 *
 * Conventions
 *  - Every function returns an int status: 0 = OK, < 0 = error. The message
 *    of the last error on a handle is returned by mocgpu_last_error().
 *  - Scalar flux arrays are [n_fsr][n_group] (group fastest), FP64, exactly
 *    the layout of TransportSweeper::flux_ (src/core/transport_sweeper.cpp:81).
 *  - Angle index iang follows mocc::AngularQuadrature: octant-major, then
 *    azimuth, then polar (src/core/product_quadrature.hpp:150-157).
 */
#ifndef MOCGPU_H
#define MOCGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (mapped from the reference's exception sites). */
#define MOCGPU_OK 0
#define MOCGPU_ERR_ARG -1        /* invalid argument / size mismatch            */
#define MOCGPU_ERR_DEVICE -2     /* HIP runtime error                           */
#define MOCGPU_ERR_NAN -3        /* keff not a number (eigen_solver.cpp:205-208) */
#define MOCGPU_ERR_NONPOS_FLUX -4/* set_pin_flux_1g <= 0 (moc_sweeper.cpp:367-375) */
#define MOCGPU_ERR_MAXITER -5    /* max_iter reached (eigen_solver.cpp:219-221)  */
#define MOCGPU_ERR_UNSUPPORTED -6/* option not implemented on the GPU path       */
#define MOCGPU_ERR_CAPACITY -7   /* request larger than an admission limit      */

/* Boundary condition kinds, same numbering as mocc::Boundary
 * (src/core/constants.hpp). */
#define MOCGPU_BC_VACUUM 0
#define MOCGPU_BC_REFLECT 1

/* Surface numbering, same as mocc::Surface (src/core/constants.hpp). */
#define MOCGPU_EAST 0
#define MOCGPU_NORTH 1
#define MOCGPU_WEST 2
#define MOCGPU_SOUTH 3
#define MOCGPU_TOP 4
#define MOCGPU_BOTTOM 5
#define MOCGPU_INVALID 7

/*
 * One 2-D MoC problem as emitted by the host ray tracer, in
 * structure-of-arrays form. All real data is FP64 (the reference's
 * real_t); the GPU library converts what its kernels read to FP32.
 *
 * Rays are shared by the polar angles of an azimuth (modularization depends
 * on the azimuth only, src/sweepers/moc/ray_data.cpp:491-527), so segment
 * data is stored once per azimuth. Azimuth a in [0, n_azi) is octant
 * (a / n_azi_oct) + 1 and azimuthal index (a % n_azi_oct) in that octant;
 * the reference angle index of (a, p) is iang = a * n_polar + p.
 */
typedef struct mocgpu_problem {
    /* --- sizes ---------------------------------------------------------- */
    int32_t n_fsr;       /* F: flat source regions                        */
    int32_t n_group;     /* G                                             */
    int32_t n_polar;     /* P: polar angles per azimuth                   */
    int32_t n_azi;       /* A: azimuths swept (octants 1 and 2)           */
    int32_t n_ray;       /* R: rays over all azimuths                     */
    int32_t n_mat;       /* XS regions (materials)                        */
    int64_t n_seg;       /* S: 2-D segments over all rays                 */

    /* --- angles: index a*P+p over octants 1-2 (2*ndir_oct entries) ---- */
    const double *ang_weight;    /* quadrature weight (sum 8 over sphere)  */
    const double *ang_sin_theta; /* sin(theta) used by sweep1g             */
    const double *ang_rsin_theta;/* 1/sin(theta) (Direction::rsintheta)    */
    const double *ang_ox;        /* direction cosines (Current weights)    */
    const double *ang_oy;
    const double *ang_alpha;     /* azimuthal angle                        */

    /* --- azimuths ------------------------------------------------------- */
    const int32_t *azi_ray_off;  /* [A+1] rays of azimuth a               */
    const int32_t *azi_nx;       /* [A] rays on the y-normal faces        */
    const int32_t *azi_ny;       /* [A] rays on the x-normal faces        */
    const double *azi_spacing;   /* [A] ray spacing after modularization  */

    /* --- rays ----------------------------------------------------------- */
    const int64_t *ray_seg_off;  /* [R+1] segments of ray r               */
    const int32_t *ray_bc;       /* [2R] entry/exit boundary index (bc(0), bc(1)) */
    const int32_t *seg_fsr;      /* [S] FSR index of each segment         */
    const double *seg_len;       /* [S] volume-corrected 2-D length       */

    /* --- coarse (pin-cell) crossings, as moc::Ray coarse data ----------- */
    const int32_t *ray_cm_cell;  /* [2R] first coarse cell fw, bw        */
    const int32_t *ray_cm_surf;  /* [2R] first coarse surface fw, bw     */
    const int32_t *ray_cm_off;   /* [R+1] entries of ray r in cm_data    */
    const int32_t *cm_data;      /* [4*n_cm] (fw surf, bw surf, nseg_fw, nseg_bw) */

    /* --- regions and cross sections -------------------------------------- */
    const double *fsr_vol;       /* [F] area * plane height               */
    const double *fsr_area;      /* [F] 2-D area (PinMesh::areas())       */
    const int32_t *fsr_mat;      /* [F] XS region of each FSR             */
    const double *xs_tr;         /* [n_mat*G] transport XS                */
    const double *xs_nf;         /* [n_mat*G] nu-fission                  */
    const double *xs_f;          /* [n_mat*G] fission                     */
    const double *xs_ch;         /* [n_mat*G] chi (normalized)            */
    const double *xs_sc;         /* [n_mat*G*G] scattering [to][from]     */

    /* --- pin (coarse) mesh ---------------------------------------------- */
    int32_t nx, ny;              /* pins along x, y                       */
    const double *x_vec;         /* [nx+1] pin boundaries                 */
    const double *y_vec;         /* [ny+1]                                */
    double hz;                   /* plane (macroplane) height             */
    const int32_t *pin_fsr_off;  /* [nx*ny+1] FSRs of pin i (FSR order)   */
    const int32_t *pin_cell;     /* [nx*ny] coarse cell of pin i          */
    const double *pin_area;      /* [nx*ny] pin mesh area (pitch_x*pitch_y) */
    int32_t bc[6];               /* mocc::Surface-indexed boundary kinds  */
    int32_t n_fuel_pins;         /* CoreMesh::n_fuel_2d()                 */
} mocgpu_problem;

/* Iteration options, the attributes of <sweeper> / <solver> / <cmfd>. */
typedef struct mocgpu_options {
    int32_t n_inner;        /* <sweeper n_inner>                           */
    int32_t energy_mode;    /* 0: Gauss-Seidel over groups (FixedSourceSolver::step),
                               1: Jacobi (all groups from the previous outer) */
    int32_t lane_map;       /* ray/direction -> lane mapping of the sweep: 2 (default)
                               length split (each wave both directions of 32 rays, a
                               band's rays longest first, directions interleaved by
                               ray parity), 1 ray-parity interleave of directions in
                               each wave over all the band's rays, 0 one direction per
                               wave, -1 autotune lane_map x band_rays at create (times
                               candidate plans on the device)                       */
    int32_t band_rays;      /* rays per band, 1..64 (0 = 64)                */
    int32_t use_cmfd;       /* <solver cmfd="t">                           */
    int32_t cmfd_max_iter;  /* <cmfd max_iter>, default 100                */
    double cmfd_k_tol;      /* <cmfd k_tol>, default 1e-6                  */
    double cmfd_psi_tol;    /* <cmfd psi_tol>, default 1e-5                */
    double cmfd_resid_reduction; /* <cmfd residual_reduction>, default 1e-3 */
    int32_t cmfd_solver;    /* 0 BiCGSTAB (Eigen default), 1 GMRES         */
    int32_t gs_boundary;    /* <sweeper boundary_update>: 1 Gauss-Seidel in angle
                               (mocc's default, moc_sweeper.cpp:105-115), 0 Jacobi */
    int32_t cmfd_negative_fixup; /* <cmfd negative_fixup>, default 0: clamp negative
                               coarse fluxes to 0 before each CMFD solve (cmfd.cpp:152-154,
                               175-184) */
    int32_t correction_worker; /* 1: the handle can run the 2-D/3-D sweeper's correction
                               sweep (mocgpu_sweep_2d3d, cmdo::MoCSweeper_2D3D); it keeps
                               a ray-major copy of the segments. Default 0 */
} mocgpu_options;

typedef struct mocgpu_handle mocgpu_handle;

/* Fill *opt with the reference defaults. */
void mocgpu_default_options(mocgpu_options *opt);

/* The options a handle runs with (band_rays and lane_map as chosen by the
 * autotuner when created with lane_map = -1). */
int mocgpu_get_options(const mocgpu_handle *h, mocgpu_options *out);

/* Build the device-resident sweeper for one problem.
 * Replaces moc::MoCSweeper::MoCSweeper (src/sweepers/moc/moc_sweeper.cpp:62-187)
 * for the data the sweep reads; ray tracing stays on the host. */
int mocgpu_create(const mocgpu_problem *prob, const mocgpu_options *opt,
                  mocgpu_handle **out);
int mocgpu_destroy(mocgpu_handle *h);

/* Give the handle another problem's cross sections: the same geometry and
 * group count (MOCGPU_ERR_ARG otherwise), new material data, and the state
 * of a new handle (mocgpu_initialize). The device buffers are reused, so a
 * server can take the next cross-section branch case of a geometry on a
 * finished case's handle without allocating or freeing device memory. No
 * reference counterpart: mocc reads its cross sections once per run. */
int mocgpu_reload_xs(mocgpu_handle *h, const mocgpu_problem *p);

/* A handle for another problem on the SAME geometry as `like` (same deck
 * geometry and ray parameters; cross sections and group count may differ, as
 * for cross-section branch cases): the ray plan, segment streams and FSR/pin
 * arrays are shared with `like` on the device, only cross sections and state
 * are new. The shared memory is released with the last handle using it. No
 * reference counterpart: mocc builds one sweeper per run. */
int mocgpu_create_like(const mocgpu_handle *like, const mocgpu_problem *p,
                       const mocgpu_options *opt, mocgpu_handle **out);

/* Device bytes held by a handle: its own state and cross sections, and the
 * geometry it may share with other handles. */
int64_t mocgpu_state_bytes(const mocgpu_handle *h);
int64_t mocgpu_geometry_bytes(const mocgpu_handle *h);
const char *mocgpu_last_error(const mocgpu_handle *h);

/* MoCSweeper::initialize (moc_sweeper.cpp:234-253): phi = 1, psi_bc = 1/(4 pi)
 * on non-vacuum faces, coarse flux = 1. */
int mocgpu_initialize(mocgpu_handle *h);

/* Copy the scalar flux [F][G] in / out (TransportSweeper::flux()). */
int mocgpu_set_flux(mocgpu_handle *h, const double *phi);
int mocgpu_get_flux(mocgpu_handle *h, double *phi);

/* Incoming boundary angular flux for all groups, [G][4*ndir_oct][face values]
 * in BoundaryCondition::data_ order (src/core/boundary_condition.cpp:33-78). */
int mocgpu_get_boundary(mocgpu_handle *h, double *bc, int64_t n);
int mocgpu_set_boundary(mocgpu_handle *h, const double *bc, int64_t n);
int64_t mocgpu_boundary_size(const mocgpu_handle *h);

/* TransportSweeper::store_old_flux (transport_sweeper.hpp:271-275). */
int mocgpu_store_old_flux(mocgpu_handle *h);

/* TransportSweeper::calc_fission_source (transport_sweeper.cpp:119-134):
 * device fs = (1/k) sum_g nuSigf phi. If fs_out != NULL it is copied out [F]. */
int mocgpu_calc_fission_source(mocgpu_handle *h, double k, double *fs_out);
/* Upload an explicit fission source [F] (FixedSourceSolver::set_fission_source). */
int mocgpu_set_fission_source(mocgpu_handle *h, const double *fs);

/* Source::initialize_group + fission + in_scatter for group g
 * (src/core/source.cpp:41-112), into the device one-group source. */
int mocgpu_build_source(mocgpu_handle *h, int32_t g);

/* MoCSweeper::sweep(g) (moc_sweeper.cpp:189-225): n_inner x {self_scatter;
 * sweep1g}; on the last inner, if with_current, the moc::Current tallies
 * (moc_current_worker.hpp:137-330). The boundary update follows
 * options.gs_boundary: Gauss-Seidel in angle by default (mocc's default,
 * moc_sweeper_kernel.inc.hpp:137-142), Jacobi in angle with gs_boundary = 0
 * (boundary_update="jacobi", moc_sweeper_kernel.inc.hpp:146-150). */
int mocgpu_sweep(mocgpu_handle *h, int32_t g, int32_t with_current);

/* One inner sweep of every group at once (sweep1g for all groups, Jacobi in
 * energy: each group's source was built beforehand), enqueued on the
 * library's stream without waiting. Used by the benchmark and batching. */
int mocgpu_sweep_all(mocgpu_handle *h, int32_t with_current);
/* Wait for all work enqueued on the handle's stream. */
int mocgpu_synchronize(mocgpu_handle *h);

/* FixedSourceSolver::step (src/solvers/fixed_source_solver.cpp:102-117) with
 * the handle's energy_mode: store old flux; for each group build the source
 * and sweep. */
int mocgpu_fss_step(mocgpu_handle *h, int32_t with_current);

/* TransportSweeper::total_fission(old) (transport_sweeper.cpp:104-117), FP64. */
int mocgpu_total_fission(mocgpu_handle *h, int32_t old, double *out);

/* TransportSweeper::flux_residual (transport_sweeper.cpp:210-223). */
int mocgpu_flux_residual(mocgpu_handle *h, double *out);

/* MoCSweeper::get_pin_flux_1g / set_pin_flux_1g with MeshTreatment::PIN_PLANE
 * (moc_sweeper.cpp:293-433). pin arrays are [nx*ny] in coarse-cell order.
 * The problems here have one plane, so MeshTreatment::PIN (one value per pin
 * and core plane, moc_sweeper.cpp:328-357, 382-392) is the same nx*ny array;
 * the other treatments (TRUE, PLANE) are rejected by the reference too
 * ("Unsupported mesh treatment", moc_sweeper.cpp:359-360, 428-430) and have no
 * entry point. g outside [0, n_group) returns MOCGPU_ERR_ARG. */
int mocgpu_get_pin_flux_1g(mocgpu_handle *h, int32_t g, double *pin);
int mocgpu_set_pin_flux_1g(mocgpu_handle *h, int32_t g, const double *pin,
                           double *resid);

/* CoarseData current / surface_flux of group g after the last sweep with
 * currents, [n_surf] each in mocc::Mesh surface order (mesh.cpp:99-147). */
int mocgpu_get_currents(mocgpu_handle *h, int32_t g, double *current,
                        double *surface_flux);
int32_t mocgpu_n_surf(const mocgpu_handle *h);

/* Write the CoarseData current / surface_flux of group g (CoarseData is a
 * plain shared struct in the reference, coarse_data.hpp:35-106); marks the
 * radial data as present for the next CMFD solve. */
int mocgpu_set_currents(mocgpu_handle *h, int32_t g, const double *current,
                        const double *surface_flux);

/* CMFD::solve (src/core/cmfd.cpp:171-279) on the device in FP64, with the
 * coarse flux taken from the current fine flux (EigenSolver::do_cmfd,
 * eigen_solver.cpp:263-291: get_pin_flux, solve, set_pin_flux). The coarse
 * mesh may hold at most 6144 pin cells (3072 in a batch, mocgpu_batch_*).
 * A non-positive coarse flux returns
 * MOCGPU_ERR_NONPOS_FLUX (MoCSweeper::set_pin_flux_1g throws there). */
int mocgpu_cmfd_step(mocgpu_handle *h, double *k);

/* MoCSweeper::update_incoming_flux (moc_sweeper.cpp:255-291): the incoming
 * angular flux at every ray's entry, for every group and angle, set from the
 * partial currents the last CMFD solve stored (2/(4 pi) (J+ + J-)) after the
 * first solve, scaled by their ratio to the previous solve's after later ones.
 * Indexed, as the reference does (moc_sweeper.hpp:217-222), by the ray's first
 * coarse cell. Needs options.use_cmfd. The reference's caller has this call
 * commented out (eigen_solver.cpp:289); the device solve loops do not make it. */
int mocgpu_update_incoming_flux(mocgpu_handle *h);

/* cmdo::MoCSweeper_2D3D::sweep(g) (src/sweepers/cmdo/moc_sweeper_2d3d.cpp:48-99)
 * for a handle created with options.correction_worker = 1 and use_cmfd = 1:
 * n_inner inner sweeps of group g (source built beforehand), the last with the
 * cmdo::CurrentCorrections worker (correction_worker.hpp:36-292,
 * correction_worker.cpp:33-160). The pin-cell Sigma_t is homogenized from the
 * flux before that sweep; the worker computes, per coarse cell, angle and
 * direction, alpha_x, alpha_y (cell-average over summed face-average angular
 * flux) and beta (flux-weighted Sigma_t over the homogenized one), and the
 * coarse currents and surface fluxes of group g with the worker's signs. */
int mocgpu_sweep_2d3d(mocgpu_handle *h, int32_t g);
/* CorrectionData (correction_data.hpp:45-100): alpha [G][4*ndir_oct][nx*ny][2]
 * (x, y normal), beta [G][4*ndir_oct][nx*ny]; 0.5 and 1 until a correction
 * sweep of the group sets them. Either pointer may be NULL. */
int mocgpu_get_corrections(mocgpu_handle *h, double *alpha, double *beta);
/* The residual triple (alpha_x, alpha_y, beta) of group g's last correction
 * sweep (CurrentCorrections::residual, correction_worker.hpp:83-91). */
int mocgpu_correction_residual(mocgpu_handle *h, int32_t g, double *out);

/* Outer (coarse power) iterations of the last CMFD solve. */
int32_t mocgpu_cmfd_iterations(const mocgpu_handle *h);
/* BiCGSTAB iterations of the last CMFD solve, summed over its power
 * iterations and groups (each group solve stops at the tolerance of
 * cmfd.cpp:202-204 or at 150 iterations, cmfd.cpp:547). */
int32_t mocgpu_cmfd_bicgstab_iterations(const mocgpu_handle *h);

/* CMFD::flux() (src/core/cmfd.hpp:76-79): the coarse flux of the last solve,
 * [n_cell][G] with n_cell = nx * ny. */
int mocgpu_cmfd_flux(mocgpu_handle *h, double *out);

/* The assembled CMFD matrix of group g from the last solve (CMFD::setup_solve,
 * cmfd.cpp:506-544) as CSR with sorted columns; returns the number of
 * entries (<= cap) or < 0. */
int mocgpu_cmfd_matrix(mocgpu_handle *h, int32_t g, int32_t *row_ptr, int32_t *col,
                       double *val, int32_t cap);

/* EigenSolver::step (eigen_solver.cpp:225-253), device-resident: optional CMFD,
 * fission source, FSS step, k update. err_psi follows eigen_solver.cpp:171-184. */
typedef struct mocgpu_eigen_state {
    double k, k_prev, err_k, err_psi;
    int32_t iterations;
    int32_t converged;
} mocgpu_eigen_state;

/* EigenSolver::solve (eigen_solver.cpp:131-223). */
int mocgpu_eigen_solve(mocgpu_handle *h, double k_tol, double psi_tol,
                       int32_t max_iter, int32_t min_iter,
                       mocgpu_eigen_state *state);
/* EigenSolver::solve recording its convergence history, as EigenSolver keeps
 * convergence_ and iteration_times_ for its output (eigen_solver.cpp:181-187,
 * 293-315): after outer iteration n, history[4n .. 4n+3] = (k, error_k,
 * error_psi, seconds since the solve began), for n < cap. The number of rows
 * written is state->iterations (at most cap). */
int mocgpu_eigen_solve_history(mocgpu_handle *h, double k_tol, double psi_tol,
                               int32_t max_iter, int32_t min_iter,
                               mocgpu_eigen_state *state, double *history, int32_t cap);
/* The start of EigenSolver::solve (eigen_solver.cpp:131-160): the sweeper's
 * initial state (mocgpu_initialize), k = 1 on the device and in *state, the
 * iteration counters cleared. mocgpu_eigen_iterate then advances the solve
 * one outer iteration at a time (the caller applies solve's convergence
 * test: err_k < k_tol, err_psi < psi_tol, iterations > min_iter). */
int mocgpu_eigen_begin(mocgpu_handle *h, mocgpu_eigen_state *state);
/* One outer iteration (EigenSolver::step + the convergence bookkeeping of solve). */
int mocgpu_eigen_iterate(mocgpu_handle *h, mocgpu_eigen_state *state);

/* ---------------------------------------------------------------------------
 * Batch engine: many independent problems advanced together. No reference
 * counterpart (mocc solves one problem per run); it serves cross-section
 * branch cases and many-lattice workloads. A batch does not own its handles,
 * but from mocgpu_batch_add until mocgpu_batch_remove (or the batch's
 * destruction) its launches own the handle's device state: single-problem
 * entry points that change a member's state return MOCGPU_ERR_ARG, reads
 * (flux, boundary fluxes, currents, total fission, residual, pin fluxes, the
 * coarse flux and matrix, corrections) first wait for the work the batch has
 * queued, a handle is in at most one batch, and mocgpu_destroy of a member
 * takes it out of its batch first.
 *
 * Every unfinished problem advances by whole outer iterations of
 * EigenSolver::step (eigen_solver.cpp:225-253) with Jacobi energy sweeps and
 * the handle's n_inner; the convergence test is EigenSolver::solve's
 * (k_tol, psi_tol, min_iter, max_iter). A handle with CMFD (options.use_cmfd)
 * runs its CMFD step at the start of every outer iteration, as do_cmfd does,
 * and tallies currents on its last inner sweep. A problem's result is bitwise
 * the same whichever other problems share the batch.
 * ------------------------------------------------------------------------- */
typedef struct mocgpu_batch mocgpu_batch;

/* budget_bytes: admission limit on the summed mocgpu_state_bytes of the
 * problems in the batch (<= 0: unlimited). use_graph: replay the outer
 * iteration as a captured HIP graph (1) or launch it eagerly (0). */
int mocgpu_batch_create(int64_t budget_bytes, int32_t use_graph, mocgpu_batch **out);
int mocgpu_batch_destroy(mocgpu_batch *b);
const char *mocgpu_batch_last_error(const mocgpu_batch *b);

/* Admit a problem: resets it to the start of EigenSolver::solve (flat flux,
 * k = 1) and returns its id. MOCGPU_ERR_CAPACITY when the budget would be
 * exceeded (nothing changes); MOCGPU_ERR_UNSUPPORTED for a handle created
 * with energy_mode != 1 (a batch sweeps Jacobi in energy only). */
int mocgpu_batch_add(mocgpu_batch *b, mocgpu_handle *h, double k_tol, double psi_tol,
                     int32_t max_iter, int32_t min_iter, int32_t *id);

/* Enqueue n_outer outer iterations of every unfinished problem (returns
 * before they complete). */
int mocgpu_batch_run(mocgpu_batch *b, int32_t n_outer);

/* Wait for the enqueued work; report problems that finished since the last
 * poll (converged, max_iter, k not a number, or a non-positive coarse flux in
 * the CMFD rescale), up to cap; state.converged is 1 only for the first.
 * Returns the count. The flux of a finished problem stays readable with
 * mocgpu_get_flux on its handle. */
int mocgpu_batch_poll(mocgpu_batch *b, int32_t *ids, mocgpu_eigen_state *states, int32_t cap);

/* Remove a problem (finished or not); its handle is the caller's again. */
int mocgpu_batch_remove(mocgpu_batch *b, int32_t id);
int32_t mocgpu_batch_size(const mocgpu_batch *b);
int64_t mocgpu_batch_bytes_used(const mocgpu_batch *b);

/* Build the band plan of a problem on the host (no device) and verify its
 * invariants: every segment once per direction in the streams (backward
 * reversed), padding, band-local FSR slots, slot lists, band orders, current
 * events. Writes a one-line report that ends with a digest of the plan's
 * arrays. No reference counterpart (a test of this build's data layout). */
int mocgpu_plan_check(const mocgpu_problem *p, int32_t band_rays, char *msg, int32_t cap);

/* Number of (2-D segment, polar, direction, group) updates per inner sweep of
 * all groups: 2 * P * S * G. */
int64_t mocgpu_updates_per_sweep(const mocgpu_handle *h);

/* Time the sweep kernel: run n_sweeps all-group inner sweeps (Jacobi energy,
 * no currents, the handle's boundary update) and report the mean time per
 * sweep of the sweep-kernel launches (both phases of the Gauss-Seidel update,
 * the single launch of the Jacobi one), measured with HIP events on the
 * library's stream, and the mean host time per sweep including the flux
 * reduction. */
int mocgpu_time_sweeps(mocgpu_handle *h, int32_t n_sweeps, double *ms_per_sweep,
                       double *kernel_ms);

/* The sweep kernel's segment attenuation 1 - exp(-tau), evaluated on the
 * current device for n host values tau[i] >= 0 into out[i]: the
 * piecewise-linear interpolant of Exponential_Linear<10000> on [-10, 0]
 * (src/core/exponential.hpp:56-112) in terms of w = 1000 tau, with nodes
 * H_j = 1 - exp(-j/1000) and slopes computed in registers, continued past
 * tau = 10 (where the reference calls exp) with the same nodes. Replaces the
 * table lookup of moc_sweeper_kernel.inc.hpp:94-99 for testing. Returns 0 or
 * MOCGPU_ERR_DEVICE. */
int32_t mocgpu_attenuation(const float *tau, float *out, int32_t n);

/* mocgpu_attenuation for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference. It calls no hipMalloc or hipFree: the
 * values are staged in a fixed per-device array of
 * 2 * MOCGPU_ATTENUATION_NOALLOC_MAX floats (2 MiB), held by each device the
 * library runs on until the process ends. n above
 * MOCGPU_ATTENUATION_NOALLOC_MAX returns MOCGPU_ERR_CAPACITY. Calls on one
 * device share that array: calls that overlap on one device race. Results
 * are bitwise those of mocgpu_attenuation. */
#define MOCGPU_ATTENUATION_NOALLOC_MAX 262144
int32_t mocgpu_attenuation_noalloc(const float *tau, float *out, int32_t n);

#ifdef __cplusplus
}
#endif
#endif /* MOCGPU_H */
