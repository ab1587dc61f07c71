// C ABI of the host front-end (include/mocc_host.h).
#include "../../../include/mocc_host.h"

#include <algorithm>
#include <cstring>
#include <string>

#include "problem.hpp"

using namespace mocc_amd;

struct mocc_host_problem {
    std::unique_ptr<Problem> prob;       // full deck
    std::unique_ptr<CoreMesh> mesh_only; // geometry-only deck
    const CoreMesh &mesh() const { return prob ? prob->mesh() : *mesh_only; }
};

namespace {
thread_local std::string g_err;

template <typename F> int guard(F &&f)
{
    try {
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
        return MOCGPU_ERR_ARG;
    }
}

void angle_row(const Angle &a, double *o)
{
    o[0] = a.alpha;
    o[1] = a.theta;
    o[2] = a.ox;
    o[3] = a.oy;
    o[4] = a.oz;
    o[5] = a.weight;
    o[6] = a.rsintheta;
}
}  // namespace

extern "C" {

const char *mocc_host_last_error(void) { return g_err.c_str(); }

int mocc_host_load(const char *xml_path, mocc_host_problem **out)
{
    return guard([&] {
        auto h = new mocc_host_problem();
        try {
            h->prob = Problem::from_file(xml_path);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return 0;
    });
}

int mocc_host_load_string(const char *xml, const char *base_dir, mocc_host_problem **out)
{
    return guard([&] {
        auto h = new mocc_host_problem();
        try {
            h->prob = Problem::from_string(xml, base_dir ? base_dir : ".");
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return 0;
    });
}

int mocc_host_load_mesh_string(const char *xml, const char *base_dir, mocc_host_problem **out)
{
    return guard([&] {
        auto doc = xml_parse_string(xml);
        auto h = new mocc_host_problem();
        try {
            h->mesh_only = std::make_unique<CoreMesh>(*doc, base_dir ? base_dir : ".");
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return 0;
    });
}

void mocc_host_free(mocc_host_problem *h) { delete h; }

const mocgpu_problem *mocc_host_view(const mocc_host_problem *h)
{
    return (h && h->prob) ? &h->prob->view() : nullptr;
}

int mocc_host_get_params(const mocc_host_problem *h, mocc_host_params *out)
{
    if (!h || !h->prob || !out) return MOCGPU_ERR_ARG;
    const SolverParams &p = h->prob->params();
    out->k_tol = p.k_tol;
    out->psi_tol = p.psi_tol;
    out->max_iter = p.max_iter;
    out->min_iter = p.min_iter;
    out->n_inner = p.n_inner;
    out->gs_boundary = p.gauss_seidel_boundary ? 1 : 0;
    out->cmfd = p.cmfd ? 1 : 0;
    out->ndir_oct = h->prob->ndir_oct();
    return 0;
}

int mocc_host_get_options(const mocc_host_problem *h, mocgpu_options *out)
{
    if (!h || !h->prob || !out) return MOCGPU_ERR_ARG;
    *out = h->prob->options();
    return 0;
}

int mocc_host_quadrature(const mocc_host_problem *h, double *out, int32_t cap)
{
    if (!h || !h->prob) return MOCGPU_ERR_ARG;
    const AngularQuadrature &q = h->prob->rays().ang_quad;
    if (q.ndir() > cap) return MOCGPU_ERR_CAPACITY;
    for (int i = 0; i < q.ndir(); i++) angle_row(q[i], out + 7 * i);
    return q.ndir();
}

int mocc_host_quadrature_xml(const char *ang_quad_xml, double *out, int32_t cap)
{
    return guard([&] {
        auto doc = xml_parse_string(ang_quad_xml);
        const XmlNode *n = doc->child("ang_quad");
        if (!n) throw Error("no <ang_quad> element");
        AngularQuadrature q(*n);
        if (q.ndir() > cap) return (int)MOCGPU_ERR_CAPACITY;
        for (int i = 0; i < q.ndir(); i++) angle_row(q[i], out + 7 * i);
        return q.ndir();
    });
}

int mocc_host_quadrature_modify(const char *ang_quad_xml, int32_t n_mod, const int32_t *idx,
                                const double *alpha, double *out, int32_t cap)
{
    return guard([&] {
        auto doc = xml_parse_string(ang_quad_xml);
        const XmlNode *n = doc->child("ang_quad");
        if (!n) throw Error("no <ang_quad> element");
        AngularQuadrature q(*n);
        for (int i = 0; i < n_mod; i++) {
            if (idx[i] < 0 || idx[i] >= q.ndir()) throw Error("angle index out of range");
            Angle a = q[idx[i]];
            a.modify_alpha(alpha[i]);
            q.modify_angle(idx[i], a);
        }
        q.update_weights();
        if (q.ndir() > cap) return (int)MOCGPU_ERR_CAPACITY;
        for (int i = 0; i < q.ndir(); i++) angle_row(q[i], out + 7 * i);
        return q.ndir();
    });
}

int mocc_host_trace_ray(const mocc_host_problem *h, double x1, double y1, double x2, double y2,
                        int32_t *seg_fsr, double *seg_len, int32_t cap, int32_t *cm_head,
                        int32_t *cm, int32_t cm_cap)
{
    if (!h) return MOCGPU_ERR_ARG;
    return guard([&] {
        Ray r(Point2(x1, y1), Point2(x2, y2), 0, 0, h->mesh());
        if (r.nseg() > cap || (int)r.cm_data.size() > cm_cap) return (int)MOCGPU_ERR_CAPACITY;
        for (int k = 0; k < r.nseg(); k++) {
            seg_fsr[k] = r.seg_index[k];
            seg_len[k] = r.seg_len[k];
        }
        cm_head[0] = r.cm_cell_fw;
        cm_head[1] = r.cm_surf_fw;
        cm_head[2] = r.cm_cell_bw;
        cm_head[3] = r.cm_surf_bw;
        cm_head[4] = (int)r.cm_data.size();
        for (size_t i = 0; i < r.cm_data.size(); i++) {
            cm[4 * i + 0] = r.cm_data[i].fw;
            cm[4 * i + 1] = r.cm_data[i].bw;
            cm[4 * i + 2] = r.cm_data[i].nseg_fw;
            cm[4 * i + 3] = r.cm_data[i].nseg_bw;
        }
        return r.nseg();
    });
}

int mocc_host_reload_xs(mocc_host_problem *h, const char *xsl_path)
{
    if (!h || !xsl_path) return MOCGPU_ERR_ARG;
    return guard([&] {
        if (!h->prob) throw Error("a geometry-only deck has no cross sections to replace");
        h->prob->reload_xs(xsl_path);
        return 0;
    });
}

int mocc_host_fsr_areas(const mocc_host_problem *h, double *out, int32_t cap)
{
    if (!h) return MOCGPU_ERR_ARG;
    VecF a = h->mesh().areas();
    if ((int)a.size() > cap) return MOCGPU_ERR_CAPACITY;
    std::memcpy(out, a.data(), a.size() * sizeof(double));
    return (int)a.size();
}

int mocc_host_get_output_info(const mocc_host_problem *h, mocc_host_output_info *out)
{
    if (!h || !h->prob || !out) return MOCGPU_ERR_ARG;
    const SolverParams &p = h->prob->params();
    const CoreMesh &m = h->prob->mesh();
    out->n_polar = p.n_polar;
    out->n_azimuthal = p.n_azimuthal;
    out->dump_fsr_flux = p.dump_fsr_flux ? 1 : 0;
    out->dump_current = p.dump_current ? 1 : 0;
    out->nx = m.nx;
    out->ny = m.ny;
    out->nz = 1;
    out->n_fuel_2d = m.n_fuel;
    out->n_group = h->prob->view().n_group;
    out->cmfd = p.cmfd ? 1 : 0;
    return 0;
}

int mocc_host_group_bounds(const mocc_host_problem *h, double *out, int32_t cap)
{
    if (!h || !h->prob) return MOCGPU_ERR_ARG;
    const VecF &b = h->prob->mesh().mat_lib.g_bounds;
    if ((int)b.size() > cap) return MOCGPU_ERR_CAPACITY;
    for (size_t i = 0; i < b.size(); i++) out[i] = b[i];
    return (int)b.size();
}

int mocc_host_pin_fuel(const mocc_host_problem *h, int32_t *out, int32_t cap)
{
    if (!h) return MOCGPU_ERR_ARG;
    const CoreMesh &m = h->mesh();
    if ((int)m.core_pins.size() > cap) return MOCGPU_ERR_CAPACITY;
    for (size_t i = 0; i < m.core_pins.size(); i++) out[i] = m.core_pins[i]->is_fuel ? 1 : 0;
    return (int)m.core_pins.size();
}

int mocc_host_coarse_surfaces(const mocc_host_problem *h, int32_t *out, int32_t cap)
{
    if (!h) return MOCGPU_ERR_ARG;
    return guard([&] {
        const CoreMesh &m = h->mesh();
        const int n = m.nx * m.ny;
        if (6 * n > cap) return (int)MOCGPU_ERR_CAPACITY;
        static const Surface order[6] = {EAST, NORTH, WEST, SOUTH, TOP, BOTTOM};
        for (int c = 0; c < n; c++)
            for (int s = 0; s < 6; s++) out[6 * c + s] = m.coarse_surf(c, order[s]);
        return n;
    });
}

// --- the mocc command line (src/input_proc.cpp:51-162) -----------------------

int32_t mocc_host_amend_xml(const char *xml, int32_t n, const char *const *amendments, char *out, int32_t cap)
{
    return guard([&] {
        if (!xml || (n > 0 && !amendments)) throw std::runtime_error("null argument");
        auto doc = xml_parse_string(xml);
        for (int32_t k = 0; k < n; k++) xml_apply_amendment(*doc, amendments[k]);
        const std::string t = xml_serialize(*doc);
        if (out && cap > 0) {
            const size_t m = std::min<size_t>(t.size(), (size_t)cap - 1);
            std::memcpy(out, t.data(), m);
            out[m] = '\0';
        }
        return (int)t.size();
    });
}

int32_t mocc_host_case_name(const char *xml, char *out, int32_t cap)
{
    return guard([&] {
        if (!xml) throw std::runtime_error("null argument");
        auto doc = xml_parse_string(xml);
        const XmlNode *c = doc->child("case_name");
        if (!c) return 0;
        if (c->text.empty()) throw std::runtime_error("<case_name> was provided, yet empty");
        if (out && cap > 0) {
            const size_t m = std::min<size_t>(c->text.size(), (size_t)cap - 1);
            std::memcpy(out, c->text.data(), m);
            out[m] = '\0';
        }
        return (int)c->text.size();
    });
}

int32_t mocc_host_dump_iterations(const mocc_host_problem *h, int32_t *out, int32_t cap)
{
    return guard([&] {
        if (!h || !h->prob) throw std::runtime_error("no solver deck loaded");
        const auto &d = h->prob->params().dump_iterations;
        for (size_t k = 0; k < d.size() && (int32_t)k < cap; k++) out[k] = d[k];
        return (int)d.size();
    });
}

}  // extern "C"
