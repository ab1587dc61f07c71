// Owner of a traced 2-D problem: the vectors behind a mocgpu_problem view
// plus the solver attributes read from the deck.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../../include/mocgpu.h"
#include "model.hpp"

namespace mocc_amd {

struct SolverParams {
    double k_tol = 1e-6, psi_tol = 1e-5;
    int max_iter = 100, min_iter = 0;
    int n_inner = 1;
    bool gauss_seidel_boundary = true;  // <sweeper boundary_update>
    bool cmfd = false;
    bool dump_fsr_flux = false;         // <sweeper dump_fsr_flux> (moc_sweeper.cpp:103)
    bool dump_current = false;          // <cmfd dump_current> (cmfd.cpp:161-163)
    int n_polar = 0, n_azimuthal = 0;   // <ang_quad> attributes, as AngularQuadrature::output writes them
    std::vector<int> dump_iterations;   // <solver><dump_iterations>, sorted (eigen_solver.cpp:84-94)
};

// Reference defaults for mocgpu_options (cmfd.cpp:62-102).
void host_default_options(mocgpu_options *o);

class Problem {
public:
    // Read an input deck (XML plus the XSL it names) and trace its rays.
    static std::unique_ptr<Problem> from_file(const std::string &xml_path);
    static std::unique_ptr<Problem> from_string(const std::string &xml, const std::string &base_dir);

    const mocgpu_problem &view() const { return view_; }
    const SolverParams &params() const { return params_; }
    mocgpu_options options() const { return opts_; }
    const CoreMesh &mesh() const { return *mesh_; }
    const RayData &rays() const { return *rays_; }
    int ndir_oct() const { return ndir_oct_; }
    // Replace the cross sections by those of another XSL file with the same
    // material names; geometry and rays are kept. The view is updated.
    void reload_xs(const std::string &xsl_path);

private:
    void build(const XmlNode &doc, const std::string &base_dir);
    void fill_view();
    void fill_xs();

    std::unique_ptr<CoreMesh> mesh_;
    std::unique_ptr<RayData> rays_;
    SolverParams params_;
    mocgpu_options opts_{};
    mocgpu_problem view_{};
    int ndir_oct_ = 0;

    std::vector<double> ang_weight_, ang_sin_, ang_rsin_, ang_ox_, ang_oy_, ang_alpha_;
    std::vector<int32_t> azi_ray_off_, azi_nx_, azi_ny_;
    std::vector<double> azi_spacing_;
    std::vector<int64_t> ray_seg_off_;
    std::vector<int32_t> ray_bc_, seg_fsr_, ray_cm_cell_, ray_cm_surf_, ray_cm_off_, cm_data_;
    std::vector<double> seg_len_;
    std::vector<double> fsr_vol_, fsr_area_;
    std::vector<int32_t> fsr_mat_;
    std::vector<double> xs_tr_, xs_nf_, xs_f_, xs_ch_, xs_sc_;
    std::vector<double> x_vec_, y_vec_, pin_area_;
    std::vector<int32_t> pin_fsr_off_, pin_cell_;
};

}  // namespace mocc_amd
