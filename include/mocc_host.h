/*
 * mocc_host.h -- C ABI of the host front-end: read a mocc input deck (XML
 * plus XSL), build the 2-D core mesh, the modularized angular quadrature and
 * the ray trace, and expose the result as a mocgpu_problem.
 *
 * Replaces the set-up half of mocc's start-up path: InputProcessor and
 * CoreMesh (src/input_proc.cpp:51-213, src/core/core_mesh.cpp:28-206), the
 * AngularQuadrature (src/core/angular_quadrature.cpp:47-113) and RayData
 * (src/sweepers/moc/ray_data.cpp:67-343). Errors are reported through int
 * status codes and mocc_host_last_error() (thread-local).
 */
#ifndef MOCC_HOST_H
#define MOCC_HOST_H

#include <stdint.h>

#include "mocgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mocc_host_problem mocc_host_problem;

typedef struct mocc_host_params {
    double k_tol, psi_tol;       /* <solver k_tol psi_tol>            */
    int32_t max_iter, min_iter;  /* <solver max_iter min_iter>        */
    int32_t n_inner;             /* <sweeper n_inner>                 */
    int32_t gs_boundary;         /* <sweeper boundary_update> (1 = gs) */
    int32_t cmfd;                /* <solver cmfd>                     */
    int32_t ndir_oct;
} mocc_host_params;

const char *mocc_host_last_error(void);

/* Read and trace a deck from a file, or from a string whose relative
 * material_lib path is resolved against base_dir. */
int mocc_host_load(const char *xml_path, mocc_host_problem **out);
int mocc_host_load_string(const char *xml, const char *base_dir, mocc_host_problem **out);
void mocc_host_free(mocc_host_problem *h);

const mocgpu_problem *mocc_host_view(const mocc_host_problem *h);
int mocc_host_get_params(const mocc_host_problem *h, mocc_host_params *out);
int mocc_host_get_options(const mocc_host_problem *h, mocgpu_options *out);

/* Modularized quadrature, all 8 octants: rows of (alpha, theta, ox, oy, oz,
 * weight, rsintheta). Returns the number of angles, or < 0. */
int mocc_host_quadrature(const mocc_host_problem *h, double *out, int32_t cap);

/* Unmodified quadrature from an <ang_quad .../> string (AngularQuadrature
 * constructor), same row layout. */
int mocc_host_quadrature_xml(const char *ang_quad_xml, double *out, int32_t cap);

/* AngularQuadrature::modify_angle of angles idx[0..n_mod) to the azimuths
 * alpha[], then update_weights() (src/core/angular_quadrature.cpp:192-352,
 * the step RayData's modularization performs); same row layout. */
int mocc_host_quadrature_modify(const char *ang_quad_xml, int32_t n_mod, const int32_t *idx,
                                const double *alpha, double *out, int32_t cap);

/* Trace one ray between two boundary points of the core of h (moc::Ray).
 * seg_* receive up to cap segments; cm_head = {cell_fw, surf_fw, cell_bw,
 * surf_bw, n_cm}; cm receives 4*n_cm ints. Returns nseg or < 0. */
int mocc_host_trace_ray(const mocc_host_problem *h, double x1, double y1, double x2,
                        double y2, int32_t *seg_fsr, double *seg_len, int32_t cap,
                        int32_t *cm_head, int32_t *cm, int32_t cm_cap);

/* Geometry-only decks (no <sweeper>): build only the core mesh, for ray tests. */
int mocc_host_load_mesh_string(const char *xml, const char *base_dir, mocc_host_problem **out);

/* Replace the cross sections of a loaded deck by those of another XSL file
 * that defines the same material names (cross-section branch cases); the
 * geometry and the rays are kept and mocc_host_view() reflects the new data.
 * Same library reader as the deck's <material_lib> (material_lib.cpp). */
int mocc_host_reload_xs(mocc_host_problem *h, const char *xsl_path);

/* Total FSR area per region of the plane, [n_fsr] (the true volumes the
 * FLAT correction preserves). */
int mocc_host_fsr_areas(const mocc_host_problem *h, double *out, int32_t cap);

/* What mocc's HDF5 output needs beyond the mocgpu_problem view. */
typedef struct mocc_host_output_info {
    int32_t n_polar, n_azimuthal; /* <ang_quad> attributes (AngularQuadrature::output,
                                     angular_quadrature.cpp:213-249)                    */
    int32_t dump_fsr_flux;        /* <sweeper dump_fsr_flux> (moc_sweeper.cpp:103)      */
    int32_t dump_current;         /* <cmfd dump_current> (cmfd.cpp:161-163)             */
    int32_t nx, ny, nz;           /* pin mesh; nz = 1 (single-plane decks only)         */
    int32_t n_fuel_2d;            /* fuel pins of the plane (CoreMesh::n_fuel_2d)       */
    int32_t n_group;
    int32_t cmfd;                 /* <solver cmfd>: CMFD::output runs only with CMFD     */
} mocc_host_output_info;
int mocc_host_get_output_info(const mocc_host_problem *h, mocc_host_output_info *out);

/* Energy-group upper bounds of the XSL library (MaterialLib::g_bounds,
 * written as "eubounds" by FixedSourceSolver::output). Returns ng. */
int mocc_host_group_bounds(const mocc_host_problem *h, double *out, int32_t cap);

/* Per pin in FSR order: 1 if the pin is fuel (Pin::is_fuel: a fissile
 * material or the pin's fuel attribute, pin.cpp:67-82). Returns the pin count. */
int mocc_host_pin_fuel(const mocc_host_problem *h, int32_t *out, int32_t cap);

/* Coarse surface index of every (cell, surface) for the six surfaces in
 * AllSurfaces order (EAST, NORTH, WEST, SOUTH, TOP, BOTTOM; constants.cpp:23)
 * as Mesh::coarse_surf numbers them: out[6 * cell + s]. Returns the cell count. */
int mocc_host_coarse_surfaces(const mocc_host_problem *h, int32_t *out, int32_t cap);

/* --- mocc's command line (InputProcessor, src/input_proc.cpp:51-162) ------- */

/* Apply n command-line amendments "path/to/name=value" (mocc -a,
 * input_proc.cpp:127-146, 217-272) to the deck text xml and write the
 * amended deck as XML text into out (NUL-terminated, truncated to cap).
 * Returns the full text length, or < 0 with mocc's message in
 * mocc_host_last_error() for a malformed amendment or a missing target. */
int32_t mocc_host_amend_xml(const char *xml, int32_t n, const char *const *amendments, char *out,
                            int32_t cap);

/* The deck's <case_name> text into out: returns its length, 0 when the deck
 * has none (mocc then names the case after the input file without its
 * extension), < 0 when the element is present but empty
 * (input_proc.cpp:149-155). */
int32_t mocc_host_case_name(const char *xml, char *out, int32_t cap);

/* <solver><dump_iterations> of the deck, sorted (eigen_solver.cpp:84-94):
 * the outer iterations after which mocc writes <case>_iter_<n>.h5. Returns
 * the count (entries beyond cap are not written). */
int32_t mocc_host_dump_iterations(const mocc_host_problem *h, int32_t *out, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif /* MOCC_HOST_H */
