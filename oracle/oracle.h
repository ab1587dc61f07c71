/*
 * oracle.h -- FP64 CPU restatement of mocc's MoC hot path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * It is test infrastructure, never linked into or called by the product.
 *
 * It consumes the same mocgpu_problem the device library receives, so both
 * sides see bit-identical segments and cross sections.
 */
#ifndef MOCC_ORACLE_H
#define MOCC_ORACLE_H

#include <stdint.h>

#include "../include/mocgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_handle oracle_handle;

typedef struct oracle_options {
    int32_t n_inner;
    int32_t gs_boundary;   /* 1: Gauss-Seidel boundary update per angle (reference default) */
    int32_t energy_mode;   /* 0: Gauss-Seidel over groups (reference), 1: Jacobi */
    int32_t exact_exp;     /* 0: Exponential_Linear<10000> table (reference), 1: std::exp */
    int32_t use_cmfd;
    int32_t cmfd_max_iter;
    double cmfd_k_tol, cmfd_psi_tol, cmfd_resid_reduction;
    int32_t n_threads;     /* 0: OpenMP default */
    int32_t cmfd_negative_fixup; /* <cmfd negative_fixup> (cmfd.cpp:152-154, 175-184) */
} oracle_options;

void oracle_default_options(oracle_options *o);
int oracle_create(const mocgpu_problem *p, const oracle_options *o, oracle_handle **out);
void oracle_destroy(oracle_handle *h);
const char *oracle_last_error(void);

int oracle_initialize(oracle_handle *h);
int oracle_set_flux(oracle_handle *h, const double *phi);
int oracle_get_flux(oracle_handle *h, double *phi);
int64_t oracle_boundary_size(const oracle_handle *h);
int oracle_get_boundary(oracle_handle *h, double *bc);
int oracle_set_boundary(oracle_handle *h, const double *bc);
int oracle_store_old_flux(oracle_handle *h);
int oracle_calc_fission_source(oracle_handle *h, double k, double *fs_out);
int oracle_set_fission_source(oracle_handle *h, const double *fs);
int oracle_build_source(oracle_handle *h, int32_t g);
int oracle_get_source(oracle_handle *h, double *src_1g, double *qbar);
int oracle_sweep(oracle_handle *h, int32_t g, int32_t with_current);
int oracle_fss_step(oracle_handle *h, int32_t with_current);
int oracle_total_fission(oracle_handle *h, int32_t old, double *out);
int oracle_flux_residual(oracle_handle *h, double *out);
int oracle_get_currents(oracle_handle *h, int32_t g, double *cur, double *sfl);
int32_t oracle_n_surf(const oracle_handle *h);
int oracle_get_pin_flux_1g(oracle_handle *h, int32_t g, double *pin);
int oracle_set_pin_flux_1g(oracle_handle *h, int32_t g, const double *pin, double *resid);
int oracle_cmfd_step(oracle_handle *h, double *k, int32_t *iters);
/* CMFD linear system of group g after setup: CSR (row_ptr[n+1], col[nnz], val[nnz]). */
int oracle_cmfd_matrix(oracle_handle *h, int32_t g, int32_t *row_ptr, int32_t *col, double *val, int32_t cap);
int oracle_cmfd_flux(oracle_handle *h, double *out);
int oracle_eigen_solve(oracle_handle *h, double k_tol, double psi_tol, int32_t max_iter,
                       int32_t min_iter, mocgpu_eigen_state *st, double *k_hist, int32_t hist_cap);
int oracle_eigen_iterate(oracle_handle *h, mocgpu_eigen_state *st);

/* MoCSweeper_2D3D::sweep(g) (moc_sweeper_2d3d.cpp:48-99): the last inner
 * sweep runs the cmdo::CurrentCorrections worker (correction_worker.hpp,
 * correction_worker.cpp:33-160), storing alpha [G][4nd][cell][2] and beta
 * [G][4nd][cell] (CorrectionData order, correction_data.hpp:45-100) and the
 * group's residual triple (alpha_x, alpha_y, beta). */
int oracle_sweep_2d3d(oracle_handle *h, int32_t g);
int oracle_get_corrections(oracle_handle *h, double *alpha, double *beta);
int oracle_correction_residual(oracle_handle *h, int32_t g, double *out);
/* MoCSweeper::update_incoming_flux (moc_sweeper.cpp:255-291). */
int oracle_update_incoming_flux(oracle_handle *h);

/* Exponential_Linear<N> on [xmin, xmax] (src/core/exponential.hpp:56-112). */
double oracle_exp_table(double v, int32_t n, double xmin, double xmax);
double oracle_exp_table_node(int32_t i, int32_t n, double xmin, double xmax);

/* Wall time of n all-group inner sweeps (sweep1g only, sources fixed), ms per
 * sweep; used as the timed CPU baseline. */
int oracle_time_sweeps(oracle_handle *h, int32_t n, double *ms_per_sweep);

/* Textbook Jacobi-preconditioned BiCGSTAB (the algorithm of Eigen's
 * BiCGSTAB<SparseMatrix, DiagonalPreconditioner>) on a CSR system. */
int oracle_bicgstab(int32_t n, const int32_t *row_ptr, const int32_t *col, const double *val,
                    const double *b, double *x, double tol, int32_t max_iter, int32_t *iters,
                    double *err);

#ifdef __cplusplus
}
#endif
#endif
