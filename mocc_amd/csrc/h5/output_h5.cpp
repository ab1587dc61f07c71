// mocc's HDF5 output file from the results of a solve (see include/mocc_h5.h
// for the dataset list and the reference functions each part follows).
#include <hdf5.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mocc_h5.h"

namespace {

thread_local std::string g_err;

struct H5Error : std::runtime_error {
    explicit H5Error(const std::string &m) : std::runtime_error(m) {}
};

// An HDF5 identifier closed by the matching H5?close when it goes.
class Id {
public:
    Id(hid_t id, herr_t (*close)(hid_t), const std::string &what) : id_(id), close_(close)
    {
        if (id_ < 0) throw H5Error(what);
    }
    Id(const Id &) = delete;
    Id &operator=(const Id &) = delete;
    ~Id()
    {
        if (id_ >= 0) close_(id_);
    }
    operator hid_t() const { return id_; }

private:
    hid_t id_;
    herr_t (*close_)(hid_t);
};

void check(herr_t e, const std::string &what)
{
    if (e < 0) throw H5Error(what);
}

// H5Node::write(path, data, dims) (h5file.cpp:105-152): NATIVE_DOUBLE.
void write_doubles(hid_t loc, const std::string &path, const double *data, const std::vector<hsize_t> &dims)
{
    Id space(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose, "dataspace for " + path);
    Id set(H5Dcreate2(loc, path.c_str(), H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           "Failed to write dataset: " + path);
    check(H5Dwrite(set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "Failed to write dataset: " + path);
}

void write_doubles(hid_t loc, const std::string &path, const std::vector<double> &v)
{
    write_doubles(loc, path, v.data(), {(hsize_t)v.size()});
}

// H5Node::write(path, int) (h5file.hpp:197-216): NATIVE_INT, one element.
void write_int(hid_t loc, const std::string &path, int v)
{
    const hsize_t one = 1;
    Id space(H5Screate_simple(1, &one, nullptr), H5Sclose, "dataspace for " + path);
    Id set(H5Dcreate2(loc, path.c_str(), H5T_NATIVE_INT, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           "Failed to write dataset: " + path);
    check(H5Dwrite(set, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "Failed to write dataset: " + path);
}

// H5Node::write(path, string) (h5file.cpp:254-267): variable-length string, scalar.
void write_string(hid_t loc, const std::string &path, const std::string &s)
{
    Id type(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
    check(H5Tset_size(type, H5T_VARIABLE), "string type");
    Id space(H5Screate(H5S_SCALAR), H5Sclose, "scalar dataspace");
    Id set(H5Dcreate2(loc, path.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           "Failed to write string data: " + path);
    const char *p = s.c_str();
    check(H5Dwrite(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p), "Failed to write string data: " + path);
}

hid_t make_group(hid_t loc, const std::string &path)
{
    const hid_t g = H5Gcreate2(loc, path.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) throw H5Error("Failed to create group '" + path + "'");
    return g;
}

std::string group_name(int g)  // "001", "002", ... as MoCSweeper::output names them
{
    char b[16];
    std::snprintf(b, sizeof b, "%03d", g + 1);
    return b;
}

template <typename F> int guard(F &&f)
{
    try {
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

void host_call(int rc, const char *what)
{
    if (rc < 0) throw H5Error(std::string(what) + ": " + mocc_host_last_error());
}

void write_file(const mocc_host_problem *h, const char *path, const mocc_output *out)
{
    if (!h || !path || !out || !out->fsr_flux) throw H5Error("null argument");
    if (out->n_iter < 0 || (out->n_iter > 0 && !out->history)) throw H5Error("bad convergence history");
    const mocgpu_problem *p = mocc_host_view(h);
    if (!p) throw H5Error("no problem view");
    mocc_host_output_info info{};
    host_call(mocc_host_get_output_info(h, &info), "output info");
    const int F = p->n_fsr, G = p->n_group, nx = info.nx, ny = info.ny, nz = info.nz;
    const int npin = nx * ny;
    std::vector<double> eub(G + 1);
    const int ng = mocc_host_group_bounds(h, eub.data(), (int)eub.size());
    host_call(ng, "group bounds");
    eub.resize(ng);
    std::vector<int32_t> fuel(npin);
    host_call(mocc_host_pin_fuel(h, fuel.data(), npin), "fuel pins");
    const double *phi = out->fsr_flux;

    // TransportSweeper::get_pin_flux (PIN treatment, one macroplane;
    // moc_sweeper.cpp:293-374): volume-weighted average over each pin's FSRs,
    // then Normalize over every pin and group (moc_sweeper.cpp:540-541).
    std::vector<double> pin(npin * (size_t)G, 0.0);  // [cell][g], Blitz (n_reg, G) row-major
    for (int g = 0; g < G; g++)
        for (int i = 0; i < npin; i++) {
            double v = 0.0, s = 0.0;
            for (int r = p->pin_fsr_off[i]; r < p->pin_fsr_off[i + 1]; r++) {
                v += p->fsr_vol[r];
                s += phi[(size_t)r * G + g] * p->fsr_vol[r];
            }
            pin[(size_t)p->pin_cell[i] * G + g] += s / v;
        }
    {
        int n = 0;
        double sum = 0.0;
        for (double v : pin) {
            if (v > 0.0) n++;
            sum += v;
        }
        const double f = (double)n / sum;
        for (double &v : pin) v *= f;
    }
    // TransportSweeper::pin_powers (transport_sweeper.cpp:139-193): per FSR
    // sum_g phi * Sigma_f * V, summed per pin, scaled to n_fuel_2d in total.
    std::vector<double> fsr_pow(F, 0.0);
    for (int r = 0; r < F; r++)
        for (int g = 0; g < G; g++)
            fsr_pow[r] += phi[(size_t)r * G + g] * p->xs_f[(size_t)p->fsr_mat[r] * G + g] * p->fsr_vol[r];
    std::vector<double> powers(npin, 0.0);
    double tot = 0.0;
    for (int i = 0; i < npin; i++)
        for (int r = p->pin_fsr_off[i]; r < p->pin_fsr_off[i + 1]; r++) {
            tot += fsr_pow[r];
            powers[p->pin_cell[i]] += fsr_pow[r];
        }
    const double scale = (double)info.n_fuel_2d / tot;
    for (double &v : powers) v *= scale;

    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Id file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose,
            std::string("Failed to open HDF5 file: ") + path);

    // EigenSolver::output: convergence (eigen_solver.cpp:293-315)
    {
        Id cg(make_group(file, "convergence"), H5Gclose, "convergence");
        const int n = out->n_iter;
        std::vector<double> k(n), ek(n), ep(n), t(n);
        for (int i = 0; i < n; i++) {
            k[i] = out->history[4 * i];
            ek[i] = out->history[4 * i + 1];
            ep[i] = out->history[4 * i + 2];
            t[i] = out->history[4 * i + 3];
        }
        const std::vector<hsize_t> dn{(hsize_t)n};
        write_doubles(cg, "k", k.data(), dn);
        write_doubles(cg, "error_k", ek.data(), dn);
        write_doubles(cg, "error_psi", ep.data(), dn);
        write_doubles(cg, "iteration_time", t);
        write_doubles(cg, "abscissae", t);
    }
    // FixedSourceSolver::output (fixed_source_solver.cpp:119-128)
    write_int(file, "ng", G);
    write_doubles(file, "eubounds", eub.data(), {(hsize_t)ng});
    // MoCSweeper::output (moc_sweeper.cpp:534-585); mesh dims reversed: nz, ny, nx
    const std::vector<hsize_t> mdims{(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
    {
        Id fg(make_group(file, "flux"), H5Gclose, "flux");
    }
    if (info.dump_fsr_flux) {
        Id sg(make_group(file, "fsr_flux"), H5Gclose, "fsr_flux");
        std::vector<double> f1(F);
        for (int g = 0; g < G; g++) {
            for (int r = 0; r < F; r++) f1[r] = phi[(size_t)r * G + g];
            write_doubles(file, "fsr_flux/" + group_name(g), f1.data(), {(hsize_t)F, 1});
        }
    }
    {
        std::vector<double> f1(npin);
        for (int g = 0; g < G; g++) {
            for (int c = 0; c < npin; c++) f1[c] = pin[(size_t)c * G + g];
            write_doubles(file, "flux/" + group_name(g), f1.data(), mdims);
        }
    }
    write_doubles(file, "pin_powers", powers.data(), mdims);
    // AngularQuadrature::output (angular_quadrature.cpp:213-249)
    {
        const int cap = 8 * 4096;
        std::vector<double> rows(7 * (size_t)cap);
        const int nd = mocc_host_quadrature(h, rows.data(), cap);
        host_call(nd, "quadrature");
        std::vector<double> col[6];
        for (int c = 0; c < 6; c++) {
            col[c].resize(nd);
            for (int i = 0; i < nd; i++) col[c][i] = rows[7 * (size_t)i + c];
        }
        Id ag(make_group(file, "ang_quad"), H5Gclose, "ang_quad");
        write_int(ag, "n_polar", info.n_polar);
        write_int(ag, "n_azimuthal", info.n_azimuthal);
        static const char *names[6] = {"alpha", "theta", "omega_x", "omega_y", "omega_z", "weight"};
        for (int c = 0; c < 6; c++) write_doubles(ag, names[c], col[c]);
    }
    // CMFD::output (cmfd.cpp:668-696), called only when CMFD is on
    // (eigen_solver.cpp:318-320)
    if (info.dump_current && info.cmfd) {
        if (!out->current || out->n_surf <= 0) throw H5Error("dump_current is set but no currents were given");
        std::vector<int32_t> surf(6 * (size_t)npin);
        host_call(mocc_host_coarse_surfaces(h, surf.data(), (int)surf.size()), "coarse surfaces");
        Id cg(make_group(file, "current"), H5Gclose, "current");
        write_string(cg, "source", "CMFD");
        static const char *sname[6] = {"east", "north", "west", "south", "top", "bottom"};
        std::vector<double> v(npin);
        for (int g = 0; g < G; g++) {
            Id gg(make_group(cg, std::to_string(g + 1)), H5Gclose, "current group");
            for (int s = 0; s < 6; s++) {
                for (int c = 0; c < npin; c++) {
                    const int k = surf[6 * (size_t)c + s];
                    if (k < 0 || k >= out->n_surf) throw H5Error("coarse surface index out of range");
                    v[c] = out->current[(size_t)g * out->n_surf + k];
                }
                write_doubles(gg, sname[s], v.data(), mdims);
            }
        }
    }
}

struct Lister {
    std::string text;
};

herr_t list_cb(hid_t g, const char *name, const H5L_info_t *, void *data)
{
    H5O_info_t oi;
    if (H5Oget_info_by_name2(g, name, &oi, H5O_INFO_BASIC, H5P_DEFAULT) < 0) return -1;
    if (oi.type == H5O_TYPE_DATASET) {
        Lister *l = (Lister *)data;
        l->text += name;
        l->text += '\n';
    }
    return 0;
}

}  // namespace

extern "C" {

const char *mocc_h5_last_error(void) { return g_err.c_str(); }

int mocc_h5_write(const mocc_host_problem *h, const char *path, const mocc_output *out)
{
    return guard([&] {
        write_file(h, path, out);
        return 0;
    });
}

int mocc_h5_write_string(const char *path, const char *name, const char *value)
{
    return guard([&] {
        if (!path || !name || !value) throw H5Error("null argument");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Id file(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose, std::string("cannot open ") + path);
        write_string(file, name, value);
        return 0;
    });
}

int64_t mocc_h5_read(const char *path, const char *name, double *out, int64_t cap, int64_t *dims, int32_t max_dims,
                     int32_t *ndim)
{
    int64_t n = -1;
    const int rc = guard([&] {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("cannot open ") + path);
        Id set(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose, std::string("no dataset ") + name);
        Id space(H5Dget_space(set), H5Sclose, "dataspace");
        const int rank = H5Sget_simple_extent_ndims(space);
        if (rank < 0) throw H5Error("dataspace rank");
        std::vector<hsize_t> d(rank > 0 ? rank : 1, 1);
        if (rank > 0) H5Sget_simple_extent_dims(space, d.data(), nullptr);
        n = 1;
        for (int i = 0; i < rank; i++) n *= (int64_t)d[i];
        if (ndim) *ndim = rank;
        for (int i = 0; i < rank && i < max_dims; i++)
            if (dims) dims[i] = (int64_t)d[i];
        if (n > cap) throw H5Error("buffer too small");
        if (n > 0) check(H5Dread(set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out), "read");
        return 0;
    });
    return rc < 0 ? -1 : n;
}

int32_t mocc_h5_read_string(const char *path, const char *name, char *out, int32_t cap)
{
    int32_t len = -1;
    const int rc = guard([&] {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("cannot open ") + path);
        Id set(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose, std::string("no dataset ") + name);
        Id type(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
        check(H5Tset_size(type, H5T_VARIABLE), "string type");
        char *s = nullptr;
        check(H5Dread(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, &s), "read string");
        const std::string v = s ? s : "";
        if (s) H5free_memory(s);
        len = (int32_t)v.size();
        if (out && cap > 0) {
            std::strncpy(out, v.c_str(), (size_t)cap - 1);
            out[cap - 1] = '\0';
        }
        return 0;
    });
    return rc < 0 ? -1 : len;
}

int32_t mocc_h5_list(const char *path, char *out, int32_t cap)
{
    int32_t len = -1;
    const int rc = guard([&] {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("cannot open ") + path);
        Lister l;
        check(H5Lvisit(file, H5_INDEX_NAME, H5_ITER_INC, list_cb, &l), "visit");
        len = (int32_t)l.text.size();
        if (out && cap > 0) {
            std::strncpy(out, l.text.c_str(), (size_t)cap - 1);
            out[cap - 1] = '\0';
        }
        return 0;
    });
    return rc < 0 ? -1 : len;
}

}  // extern "C"
