// mocc's command-line program for the gfx950 path:
//
//     mocc [-a path/to/attribute=value ...] infile
//
// Restates the reference's driver (src/mocc.cpp, src/driver.cpp:51-182) and
// InputProcessor (src/input_proc.cpp:51-162) over this build's C ABIs:
//   * libmocc_host: the deck and its amendments, the core mesh, the ray trace;
//   * libmocgpu:    the eigenvalue solve on the device (sweeps, sources, CMFD),
//                   stepped one outer iteration at a time as EigenSolver::solve
//                   steps it (eigen_solver.cpp:131-223);
//   * libmocc_h5:   the output file <case_name>.h5 and the snapshots
//                   <case_name>_iter_<n>.h5 of <dump_iterations>, loaded at run
//                   time after the HDF5 C library (its directory also holds an
//                   older libstdc++, so it is not on the run path).
// The log goes to <case_name>.log and, for what mocc prints to the screen,
// also to stdout. The case name is the input file name without its extension
// unless the deck has a <case_name>.
//
// SIGINT: mocc's handler writes the output file and exits with failure
// (driver.cpp:93-99). Here the handler only raises a flag; the solve loop
// checks it after every outer iteration, then writes the output and exits
// with failure (writing a file from inside a signal handler is not safe).
#include <dlfcn.h>
#include <unistd.h>

#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mocc_h5.h"
#include "../../../include/mocc_host.h"
#include "../../../include/mocgpu.h"

#ifndef MOCC_GIT_SHA1
#define MOCC_GIT_SHA1 "unknown"
#endif

namespace {

volatile std::sig_atomic_t g_interrupted = 0;
void on_sigint(int) { g_interrupted = 1; }

// Writes to the log file and, for screen output, to stdout as well
// (mocc's LogFile / LogScreen, util/files.hpp).
struct Log {
    std::ofstream file;
    void screen(const std::string &s)
    {
        std::cout << s << std::flush;
        if (file) file << s << std::flush;
    }
    void log(const std::string &s)
    {
        if (file) file << s << std::flush;
    }
};

struct Usage : std::runtime_error {
    explicit Usage(const std::string &m) : std::runtime_error(m) {}
};

// InputProcessor's command line (input_proc.cpp:51-103).
void parse_args(int argc, char **argv, std::string &filename, std::vector<std::string> &amendments)
{
    for (int i = 1; i < argc; i++) {
        const std::string arg = argv[i];
        if (arg == "-a") {
            if (i == argc - 1 || argv[i + 1][0] == '-') throw Usage("-a option specified without argument");
            amendments.push_back(argv[++i]);
        } else if (filename.empty()) {
            filename = arg;
        } else {
            throw Usage("Filename appears to be multiply-defined: " + filename);
        }
    }
    if (filename.empty()) throw Usage("No filename");
}

std::string strip_extension(const std::string &s) { return s.substr(0, s.rfind('.')); }

std::string dir_of(const std::string &path)
{
    const size_t p = path.rfind('/');
    return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

std::string read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("XML parse error: File was not found\nError encountered in parsing XML file.");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void check_host(int rc)
{
    if (rc < 0) throw std::runtime_error(mocc_host_last_error());
}

// The output library, after libhdf5 by its full path.
struct H5 {
    decltype(&mocc_h5_write) write = nullptr;
    decltype(&mocc_h5_write_string) write_string = nullptr;
    decltype(&mocc_h5_last_error) last_error = nullptr;

    void load()
    {
        const char *pre = std::getenv("MOCC_HDF5_PREFIX");
        const std::string hdf5 = std::string(pre ? pre : "/opt/conda") + "/lib/libhdf5.so.103";
        if (!dlopen(hdf5.c_str(), RTLD_NOW | RTLD_GLOBAL))
            throw std::runtime_error("the HDF5 C library is not at " + hdf5 + " (set MOCC_HDF5_PREFIX)");
        char exe[4096];
        const ssize_t n = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
        if (n <= 0) throw std::runtime_error("cannot locate the executable");
        exe[n] = '\0';
        const std::string lib = dir_of(exe) + "/libmocc_h5.so";
        void *h = dlopen(lib.c_str(), RTLD_NOW);
        if (!h) throw std::runtime_error(std::string("cannot load ") + lib + ": " + dlerror());
        write = (decltype(write))dlsym(h, "mocc_h5_write");
        write_string = (decltype(write_string))dlsym(h, "mocc_h5_write_string");
        last_error = (decltype(last_error))dlsym(h, "mocc_h5_last_error");
        if (!write || !write_string || !last_error) throw std::runtime_error("libmocc_h5.so lacks its entry points");
    }
    void check(int rc) const
    {
        if (rc < 0) throw std::runtime_error(last_error());
    }
};

struct Run {
    Log log;
    H5 h5;
    std::string case_name, deck_text;
    mocc_host_problem *hp = nullptr;
    mocgpu_handle *dev = nullptr;
    mocc_host_output_info info{};
    std::vector<double> history;  // (k, error_k, error_psi, seconds) per outer iteration
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    ~Run()
    {
        if (dev) mocgpu_destroy(dev);
        if (hp) mocc_host_free(hp);
    }

    double seconds() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void check_dev(int rc) const
    {
        if (rc != 0) throw std::runtime_error(mocgpu_last_error(dev));
    }

    // EigenSolver::output (eigen_solver.cpp:293-322) of the state so far.
    void write_output(const std::string &path, bool with_input)
    {
        const mocgpu_problem *p = mocc_host_view(hp);
        const int F = p->n_fsr, G = p->n_group;
        std::vector<double> phi((size_t)F * G);
        check_dev(mocgpu_get_flux(dev, phi.data()));
        std::vector<double> cur;
        mocc_output out{};
        out.fsr_flux = phi.data();
        out.n_iter = (int32_t)(history.size() / 4);
        out.history = history.empty() ? nullptr : history.data();
        if (info.cmfd && info.dump_current) {
            const int ns = mocgpu_n_surf(dev);
            cur.assign((size_t)G * ns, 0.0);
            for (int g = 0; g < G; g++) check_dev(mocgpu_get_currents(dev, g, cur.data() + (size_t)g * ns, nullptr));
            out.current = cur.data();
            out.n_surf = ns;
        }
        h5.check(h5.write(hp, path.c_str(), &out));
        if (with_input) {
            // driver.cpp:59-68: the deck as run (with its amendments) and the build
            h5.check(h5.write_string(path.c_str(), "input_file", deck_text.c_str()));
            h5.check(h5.write_string(path.c_str(), "git_sha1", MOCC_GIT_SHA1));
        }
    }

    void generate_output()
    {
        write_output(case_name + ".h5", true);
        std::cout << "Output written to '" << case_name << "'" << std::endl;
    }

    // EigenSolver::solve (eigen_solver.cpp:131-223) with the device step.
    void solve()
    {
        mocc_host_params par{};
        check_host(mocc_host_get_params(hp, &par));
        std::vector<int32_t> dumps(64);
        int nd = mocc_host_dump_iterations(hp, dumps.data(), (int32_t)dumps.size());
        check_host(nd);
        if (nd > (int)dumps.size()) {
            dumps.resize(nd);
            check_host(mocc_host_dump_iterations(hp, dumps.data(), nd));
        }
        dumps.resize(nd);
        if (nd > 0) {
            std::ostringstream m;
            m << "Dumping data for eigenvalue iterations:\n";
            for (int d : dumps) m << d << " ";
            log.log(m.str() + "\n");
        }
        size_t next = 0;
        mocgpu_eigen_state st{};
        check_dev(mocgpu_eigen_begin(dev, &st));
        for (int n = 0; n < par.max_iter; n++) {
            const int rc = mocgpu_eigen_iterate(dev, &st);
            if (rc == MOCGPU_ERR_NAN) throw std::runtime_error("Eigenvalue is not a number. Giving up.");
            check_dev(rc);
            history.insert(history.end(), {st.k, st.err_k, st.err_psi, seconds()});
            std::ostringstream line;
            line << std::setw(14) << std::fixed << std::setprecision(5) << seconds() << std::setw(14) << n + 1
                 << std::setw(14) << std::fixed << std::setprecision(10) << st.k << std::setw(14)
                 << std::scientific << std::setprecision(6) << st.err_k << std::setw(14) << st.err_psi << "\n";
            log.screen(line.str());
            while (next < dumps.size() && dumps[next] < n + 1) next++;
            if (next < dumps.size() && dumps[next] == n + 1) {
                write_output(case_name + "_iter_" + std::to_string(n + 1) + ".h5", false);
                next++;
            }
            if (g_interrupted) {
                std::cout << "Caught SIGINT. Bailing." << std::endl;
                generate_output();
                std::exit(EXIT_FAILURE);
            }
            if (st.err_k < par.k_tol && st.err_psi < par.psi_tol && n >= par.min_iter) {
                log.screen("Convergence criteria satisfied!\n");
                break;
            }
            if (n == par.max_iter - 1) log.screen("Maximum number of iterations reached!\n");
        }
    }
};

int run(int argc, char **argv)
{
    std::signal(SIGINT, on_sigint);
    std::cout << "mocc (gfx950 MoC sweeper, source and CMFD on the device)" << std::endl;
    std::string filename;
    std::vector<std::string> amendments;
    try {
        parse_args(argc, argv, filename, amendments);
    } catch (const Usage &u) {
        std::cerr << u.what() << std::endl;
        std::cout << "Usage: mocc [-a substitution/path/attribute=value] infile" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    Run r;
    try {
        r.case_name = strip_extension(filename);
        std::cout << "Parsing: " << filename << std::endl;
        const std::string raw = read_file(filename);
        std::vector<const char *> am;
        for (const auto &a : amendments) {
            std::cout << "amendment " << a.substr(0, a.find('=')) << "\n";
            am.push_back(a.c_str());
        }
        const int len = mocc_host_amend_xml(raw.c_str(), (int32_t)am.size(), am.data(), nullptr, 0);
        check_host(len);
        std::vector<char> buf((size_t)len + 1);
        check_host(mocc_host_amend_xml(raw.c_str(), (int32_t)am.size(), am.data(), buf.data(), len + 1));
        r.deck_text.assign(buf.data(), (size_t)len);
        const int cn = mocc_host_case_name(r.deck_text.c_str(), nullptr, 0);
        check_host(cn);
        if (cn > 0) {
            std::vector<char> c((size_t)cn + 1);
            check_host(mocc_host_case_name(r.deck_text.c_str(), c.data(), cn + 1));
            r.case_name.assign(c.data(), (size_t)cn);
        }
        const std::string logname = r.case_name + ".log";
        std::cout << "Logging output to: " << logname << std::endl << std::endl;
        r.log.file.open(logname, std::ios::out);
        r.log.screen("Running case: " + r.case_name + "\n");
        r.log.screen(std::string("Using MOCC executable built with GIT SHA1: ") + MOCC_GIT_SHA1 + "\n");
        {
            std::time_t t = std::time(nullptr);
            r.log.screen(std::string("Local time: ") + std::ctime(&t) + "\n\n");
        }
        r.log.log("XML input (including command-line amendments):\n" + std::string(79, '=') + "\n" +
                  r.deck_text + std::string(79, '=') + "\n\n");
        r.h5.load();
        check_host(mocc_host_load_string(r.deck_text.c_str(), dir_of(filename).c_str(), &r.hp));
        check_host(mocc_host_get_output_info(r.hp, &r.info));
        mocgpu_options opt;
        check_host(mocc_host_get_options(r.hp, &opt));
        if (mocgpu_create(mocc_host_view(r.hp), &opt, &r.dev) != 0)
            throw std::runtime_error(mocgpu_last_error(nullptr));
        {
            const mocgpu_problem *p = mocc_host_view(r.hp);
            std::ostringstream m;
            m << "Problem: " << p->n_fsr << " flat source regions, " << p->n_group << " groups, " << p->n_seg
              << " segments; device solve\n";
            r.log.screen(m.str());
        }
        r.solve();
        r.generate_output();
        std::ostringstream m;
        m << "Total time: " << std::fixed << std::setprecision(5) << r.seconds() << " s\n";
        r.log.screen(m.str());
    } catch (const std::exception &e) {
        std::cerr << "Error:" << std::endl << e.what() << std::endl;
        r.log.log(std::string("Error:\n") + e.what() + "\n");
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) { return run(argc, argv); }
