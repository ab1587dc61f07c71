/* mocc's solver output file (HDF5) for a problem solved through mocgpu.
 *
 * mocc writes one file per run from EigenSolver::output
 * (src/driver.cpp:51-57 -> src/solvers/eigen_solver.cpp:293-322), which
 * calls FixedSourceSolver::output (fixed_source_solver.cpp:119-128),
 * MoCSweeper::output (src/sweepers/moc/moc_sweeper.cpp:534-585),
 * TransportSweeper::pin_powers (src/core/transport_sweeper.cpp:139-193),
 * AngularQuadrature::output (src/core/angular_quadrature.cpp:213-249) and,
 * with CMFD on and <cmfd dump_current="true">, CMFD::output
 * (src/core/cmfd.cpp:668-696).
 * mocc_h5_write produces the same datasets (names, shapes, types) from the
 * FSR flux, the convergence history and the coarse currents of a solve:
 *
 *   /ng                        int [1]
 *   /eubounds                  double [ng]
 *   /flux/001 ..               double [nz][ny][nx] pin flux, Normalize'd over
 *                              all pins and groups (utils.hpp:28-51)
 *   /fsr_flux/001 ..           double [F][1] (only with dump_fsr_flux)
 *   /pin_powers                double [nz][ny][nx], sum over pins of
 *                              phi * Sigma_f * V scaled to the fuel-pin count
 *   /ang_quad/n_polar, n_azimuthal          int [1]
 *   /ang_quad/alpha, theta, omega_x, omega_y, omega_z, weight   double [ndir]
 *   /convergence/k, error_k, error_psi      double [n_iter]
 *   /convergence/iteration_time, abscissae  double [n_iter]
 *   /current/source            variable-length string "CMFD"
 *   /current/<g>/east .. bottom  double [nz][ny][nx] (only with CMFD on and
 *                              dump_current)
 *
 * The library is separate from libmocc_host and libmocgpu so that only the
 * output path loads libhdf5. Functions return 0 (or a count) on success and
 * < 0 on failure; mocc_h5_last_error() gives the message. */
#ifndef MOCC_H5_H
#define MOCC_H5_H

#include <stdint.h>

#include "mocc_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mocc_output {
    const double *fsr_flux;  /* [F * G], group fastest (TransportSweeper::flux_)       */
    int32_t n_iter;          /* rows of history                                        */
    const double *history;   /* [4 * n_iter]: k, error_k, error_psi, seconds           */
    const double *current;   /* [G * n_surf] coarse-surface currents, or NULL           */
    int32_t n_surf;          /* surfaces per group in current (mocgpu_n_surf)           */
} mocc_output;

const char *mocc_h5_last_error(void);

/* Write the output file of the problem h (truncating an existing file). */
int mocc_h5_write(const mocc_host_problem *h, const char *path, const mocc_output *out);

/* Add a scalar variable-length string dataset to an existing file, as the
 * driver adds "input_file" (the deck with its amendments) and "git_sha1"
 * (src/driver.cpp:59-68). */
int mocc_h5_write_string(const char *path, const char *name, const char *value);

/* Read a numeric dataset (converted to double) into out[0 .. cap). dims
 * receives up to max_dims extents and ndim the rank. Returns the element
 * count. */
int64_t mocc_h5_read(const char *path, const char *name, double *out, int64_t cap, int64_t *dims,
                     int32_t max_dims, int32_t *ndim);

/* Read a scalar variable-length string dataset into out (NUL-terminated).
 * Returns its length. */
int32_t mocc_h5_read_string(const char *path, const char *name, char *out, int32_t cap);

/* The paths of every dataset in the file, one per line, in name order.
 * Returns the number of bytes needed (excluding the NUL). */
int32_t mocc_h5_list(const char *path, char *out, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif /* MOCC_H5_H */
