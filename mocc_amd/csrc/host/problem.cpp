#include "problem.hpp"

#include <sstream>

#include <algorithm>
#include <map>

namespace mocc_amd {

void host_default_options(mocgpu_options *o)
{
    *o = mocgpu_options{};
    o->n_inner = 1;
    o->energy_mode = 0;
    o->lane_map = 2;
    o->band_rays = 0;
    o->use_cmfd = 0;
    o->cmfd_max_iter = 100;
    o->cmfd_k_tol = 1.0e-6;
    o->cmfd_psi_tol = 1.0e-5;
    o->cmfd_resid_reduction = 0.001;
    o->cmfd_solver = 0;
    o->gs_boundary = 1;
    o->cmfd_negative_fixup = 0;
    o->correction_worker = 0;
}

namespace {
std::string dirname_of(const std::string &p)
{
    size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

// TransportSweeper::find_angquad (transport_sweeper.cpp:30-55): look for
// <ang_quad> in the node, then its parents.
const XmlNode *find_angquad(const XmlNode *n)
{
    while (n) {
        if (const XmlNode *q = n->child("ang_quad")) return q;
        n = n->parent;
    }
    throw Error("Reached document root without finding an angular quadrature specification.");
}
}  // namespace

std::unique_ptr<Problem> Problem::from_file(const std::string &xml_path)
{
    auto doc = xml_parse_file(xml_path);
    auto p = std::unique_ptr<Problem>(new Problem());
    p->build(*doc, dirname_of(xml_path));
    return p;
}

std::unique_ptr<Problem> Problem::from_string(const std::string &xml, const std::string &base_dir)
{
    auto doc = xml_parse_string(xml);
    auto p = std::unique_ptr<Problem>(new Problem());
    p->build(*doc, base_dir);
    return p;
}

void Problem::build(const XmlNode &doc, const std::string &base_dir)
{
    mesh_ = std::make_unique<CoreMesh>(doc, base_dir);
    host_default_options(&opts_);

    // <solver> / <sweeper> (eigen_solver.cpp:36-126, moc_sweeper.cpp:62-187)
    const XmlNode *solver = doc.child("solver");
    const XmlNode *sweeper = solver ? solver->child("sweeper") : doc.child("sweeper");
    if (!sweeper) throw Error("No <sweeper> specified.");
    if (sweeper->attr_str("type", "moc") != "moc") throw Error("Only the moc sweeper is supported");
    if (solver) {
        params_.k_tol = solver->attr_double("k_tol", 1e-6);
        params_.psi_tol = solver->attr_double("psi_tol", 1e-5);
        params_.max_iter = solver->attr_int("max_iter", 100);
        params_.min_iter = solver->attr_int("min_iter", 0);
        params_.cmfd = solver->attr_bool("cmfd", false);
        if (const XmlNode *d = solver->child("dump_iterations")) {
            // explode_string<int> of the element's text, then sorted
            std::istringstream in(d->text);
            int v;
            while (in >> v) params_.dump_iterations.push_back(v);
            std::sort(params_.dump_iterations.begin(), params_.dump_iterations.end());
        }
        if (const XmlNode *c = solver->child("cmfd")) {
            opts_.cmfd_k_tol = c->attr_double("k_tol", opts_.cmfd_k_tol);
            opts_.cmfd_psi_tol = c->attr_double("psi_tol", opts_.cmfd_psi_tol);
            opts_.cmfd_resid_reduction = c->attr_double("residual_reduction", opts_.cmfd_resid_reduction);
            opts_.cmfd_max_iter = c->attr_int("max_iter", opts_.cmfd_max_iter);
            opts_.cmfd_negative_fixup = c->attr_bool("negative_fixup", false) ? 1 : 0;
            if (c->has_attr("enabled") && !c->attr_bool("enabled", true)) params_.cmfd = false;
            params_.dump_current = c->attr_bool("dump_current", false);
        }
    }
    int n_inner = sweeper->attr_int("n_inner", -1);
    if (n_inner < 0) throw Error("Invalid number of inner iterations specified (n_inner).");
    params_.n_inner = n_inner;
    opts_.n_inner = n_inner;
    opts_.use_cmfd = params_.cmfd ? 1 : 0;
    if (sweeper->has_attr("boundary_update")) {
        std::string b = sweeper->attr_str("boundary_update");
        std::transform(b.begin(), b.end(), b.begin(), ::tolower);
        if (b == "jacobi" || b == "j") params_.gauss_seidel_boundary = false;
        else if (b == "gs") params_.gauss_seidel_boundary = true;
        else throw Error("Unrecognized boundary update option.");
    }
    opts_.gs_boundary = params_.gauss_seidel_boundary ? 1 : 0;
    params_.dump_fsr_flux = sweeper->attr_bool("dump_fsr_flux", false);
    const XmlNode *rays = sweeper->child("rays");
    if (!rays) throw Error("No input privided for ray spacing.");
    AngularQuadrature quad(*find_angquad(sweeper));
    ndir_oct_ = quad.ndir_oct();
    params_.n_polar = quad.n_polar();
    params_.n_azimuthal = quad.n_azimuthal();
    rays_ = std::make_unique<RayData>(*rays, quad, *mesh_);
    fill_view();
}

void Problem::fill_view()
{
    const CoreMesh &m = *mesh_;
    const RayData &rd = *rays_;
    const AngularQuadrature &q = rd.ang_quad;
    int nd = q.ndir_oct();
    int A = (int)rd.rays.size();
    int P = 2 * nd / A;

    for (int i = 0; i < 2 * nd; i++) {
        const Angle &a = q[i];
        ang_weight_.push_back(a.weight);
        ang_sin_.push_back(std::sin(a.theta));
        ang_rsin_.push_back(a.rsintheta);
        ang_ox_.push_back(a.ox);
        ang_oy_.push_back(a.oy);
        ang_alpha_.push_back(a.alpha);
    }
    // per-ray records and offsets first, then the segments and coarse data of
    // each azimuth copied to their places in parallel
    azi_ray_off_.push_back(0);
    ray_seg_off_.push_back(0);
    ray_cm_off_.push_back(0);
    for (int a = 0; a < A; a++) {
        int iang = a * P;
        azi_nx_.push_back(rd.Nx[iang]);
        azi_ny_.push_back(rd.Ny[iang]);
        azi_spacing_.push_back(rd.spacing[iang]);
        for (const Ray &r : rd.rays[a]) {
            ray_seg_off_.push_back(ray_seg_off_.back() + r.nseg());
            ray_bc_.push_back(r.bc[0]);
            ray_bc_.push_back(r.bc[1]);
            ray_cm_cell_.push_back(r.cm_cell_fw);
            ray_cm_cell_.push_back(r.cm_cell_bw);
            ray_cm_surf_.push_back(r.cm_surf_fw);
            ray_cm_surf_.push_back(r.cm_surf_bw);
            ray_cm_off_.push_back(ray_cm_off_.back() + (int32_t)r.cm_data.size());
        }
        azi_ray_off_.push_back((int32_t)(ray_seg_off_.size() - 1));
    }
    seg_fsr_.resize((size_t)ray_seg_off_.back());
    seg_len_.resize((size_t)ray_seg_off_.back());
    cm_data_.resize((size_t)ray_cm_off_.back() * 4);
    parallel_for(A, [&](int a) {
        int ir = azi_ray_off_[a];
        for (const Ray &r : rd.rays[a]) {
            const int64_t s0 = ray_seg_off_[ir];
            for (int k = 0; k < r.nseg(); k++) {
                seg_fsr_[s0 + k] = r.seg_index[k];
                seg_len_[s0 + k] = r.seg_len[k];
            }
            size_t c0 = (size_t)ray_cm_off_[ir] * 4;
            for (const auto &c : r.cm_data) {
                cm_data_[c0++] = c.fw;
                cm_data_[c0++] = c.bw;
                cm_data_[c0++] = c.nseg_fw;
                cm_data_[c0++] = c.nseg_bw;
            }
            ir++;
        }
    });

    fill_xs();
    VecF areas = m.areas();
    fsr_area_ = areas;
    for (double a : areas) fsr_vol_.push_back(a * m.hz);

    x_vec_ = m.x_vec;
    y_vec_ = m.y_vec;
    pin_fsr_off_.push_back(0);
    for (size_t ip = 0; ip < m.core_pins.size(); ip++) {
        const Pin *pin = m.core_pins[ip];
        pin_fsr_off_.push_back(pin_fsr_off_.back() + pin->mesh->n_reg);
        pin_cell_.push_back(m.coarse_cell(m.pin_position_x((int)ip), m.pin_position_y((int)ip)));
        pin_area_.push_back(pin->mesh->pitch_x * pin->mesh->pitch_y);
    }

    mocgpu_problem &v = view_;
    v.n_fsr = m.n_reg;
    v.n_polar = P;
    v.n_azi = A;
    v.n_ray = (int32_t)(ray_seg_off_.size() - 1);
    v.n_seg = (int64_t)seg_fsr_.size();
    v.ang_weight = ang_weight_.data();
    v.ang_sin_theta = ang_sin_.data();
    v.ang_rsin_theta = ang_rsin_.data();
    v.ang_ox = ang_ox_.data();
    v.ang_oy = ang_oy_.data();
    v.ang_alpha = ang_alpha_.data();
    v.azi_ray_off = azi_ray_off_.data();
    v.azi_nx = azi_nx_.data();
    v.azi_ny = azi_ny_.data();
    v.azi_spacing = azi_spacing_.data();
    v.ray_seg_off = ray_seg_off_.data();
    v.ray_bc = ray_bc_.data();
    v.seg_fsr = seg_fsr_.data();
    v.seg_len = seg_len_.data();
    v.ray_cm_cell = ray_cm_cell_.data();
    v.ray_cm_surf = ray_cm_surf_.data();
    v.ray_cm_off = ray_cm_off_.data();
    v.cm_data = cm_data_.data();
    v.fsr_vol = fsr_vol_.data();
    v.fsr_area = fsr_area_.data();
    v.nx = m.nx;
    v.ny = m.ny;
    v.x_vec = x_vec_.data();
    v.y_vec = y_vec_.data();
    v.hz = m.hz;
    v.pin_fsr_off = pin_fsr_off_.data();
    v.pin_cell = pin_cell_.data();
    v.pin_area = pin_area_.data();
    for (int i = 0; i < 6; i++) v.bc[i] = (int32_t)m.bc[i];
    v.n_fuel_pins = m.n_fuel;
}

// XS mesh (xs_mesh.cpp:27-118): one XS region per material, ordered by
// material ID; fsr_mat maps every FSR to its region.
void Problem::fill_xs()
{
    const CoreMesh &m = *mesh_;
    const MaterialLib &lib = m.mat_lib;
    const int G = lib.n_group;
    std::map<int, std::vector<int>> fsr_map;
    {
        int ireg = 0;
        for (const Pin *pin : m.core_pins) {
            int ixs = 0;
            for (int mid : pin->mat_ids) {
                for (int r = 0; r < pin->mesh->n_fsrs(ixs); r++) fsr_map[mid].push_back(ireg++);
                ixs++;
            }
        }
    }
    fsr_mat_.assign(m.n_reg, -1);
    xs_tr_.clear();
    xs_nf_.clear();
    xs_f_.clear();
    xs_ch_.clear();
    xs_sc_.clear();
    int imat = 0;
    for (const auto &kv : fsr_map) {
        const Material &mat = lib.by_id(kv.first);
        for (int g = 0; g < G; g++) {
            xs_tr_.push_back(mat.xstr[g]);
            xs_nf_.push_back(mat.xsnf[g]);
            xs_f_.push_back(mat.xsf[g]);
            xs_ch_.push_back(mat.xsch[g]);
        }
        for (int to = 0; to < G; to++)
            for (int from = 0; from < G; from++) xs_sc_.push_back(mat.scat[to][from]);
        for (int f : kv.second) fsr_mat_[f] = imat;
        imat++;
    }
    mocgpu_problem &v = view_;
    v.n_group = G;
    v.n_mat = imat;
    v.fsr_mat = fsr_mat_.data();
    v.xs_tr = xs_tr_.data();
    v.xs_nf = xs_nf_.data();
    v.xs_f = xs_f_.data();
    v.xs_ch = xs_ch_.data();
    v.xs_sc = xs_sc_.data();
}

void Problem::reload_xs(const std::string &xsl_path)
{
    mesh_->mat_lib.reload(xsl_path);
    fill_xs();
}

}  // namespace mocc_amd
