// fk_mc_gpu -- the reference's Monte Carlo program (prog/fk_mc_exec.cpp with
// include/fk_mc/fk_mc.hxx) run as a batch of chains on one GPU.
//
// Parameters keep the reference's names and defaults (fk_mc.hxx:176-207,
// src/mc_metropolis.cpp:11-19, prog/fk_mc_exec.cpp:182-219) and are read as the
// reference's alps::params reads them: an optional ini file (`name = value` lines) as
// the first positional argument, then `--name value` or `--name=value`.  The
// reference runs one chain per MPI rank seeded SEED + rank; here `--chains R` runs R
// such chains at once (chain r is rank r), and `--lattice` picks the lattice the
// reference compiles into separate executables (fk_mc_cubic1d, _cubic2d, _cubic3d,
// _triangular).
//
// The run follows fk_mc::initialize and mcbase::run: randomize_f(Nf_start) from each
// chain's generator, the moves flip / add_remove / reshuffle registered in the
// reference's order with their weights, ntherm_sweeps unmeasured sweeps, then nsweeps
// sweeps of sweep_len moves each followed by the measures (energy, spectrum, spectrum
// and f-occupation histories when measure_history, IPR when measure_ipr, stiffness and
// conductivity when measure_stiffness, eigenvectors when measure_eigenfunctions), all kept
// on the device until the end.
//
// Output (`output`: an HDF5 file when it ends in .h5, as the reference writes, else a
// directory of .npy files with the same paths): the observables of observables_t
// (include/fk_mc/fk_mc.hpp:13-28) under mc_data/ with the chain as the leading axis (dropped
// in an HDF5 file of one chain, which then has the reference's shapes) -- energies,
// d2energies, c_energies [chains][nsweeps], spectrum [chains][N] (running mean), spectrum_history, focc_history, ipr_history
// [chains][nsweeps][N], stiffness [chains][nsweeps] and cond_history [chains][nsweeps][nw]
// with measure_stiffness on the reference's conductivity grid, eig_history
// [chains][nsweeps][N][N] with save_eigenfunctions -- plus, with measure_eigenfunctions, the
// outputs of save_gwr under stats/ (tdos_gwr, tdos<w>, gr_full_w<w>_re/_im, gr_w<w>, gk_w<w>
// at w = 0 + i dos_offset) -- plus dos (the DOS of the recorded spectra on the reference's
// real grid, prog/data_save.hxx:67-69,264-345), the parameters (/parameters in an HDF5
// file, parameters.txt in a directory) and stats.txt (means with errors from the spread over chains, and the specific heat
// beta^2 (<E^2> - <d2E> - <E>^2) / N of prog/data_save.hxx:185-190), and the reference's
// binning / jackknife analysis under binning/ and stats/ (energies, cv, dos0 / dos_err and,
// with measure_ipr, ipr0 / ipr_err).
#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "binning.hpp"
#include "fkgpu.h"

namespace {

struct param {
    std::string value, help;
};

struct params {
    std::vector<std::string> order;
    std::map<std::string, param> p;
    void define(const std::string& k, const std::string& v, const std::string& h) {
        order.push_back(k);
        p[k] = {v, h};
    }
    bool has(const std::string& k) const { return p.count(k) != 0; }
    const std::string& str(const std::string& k) const { return p.at(k).value; }
    double d(const std::string& k) const { return std::stod(str(k)); }
    long i(const std::string& k) const { return std::stol(str(k)); }
    bool b(const std::string& k) const {
        const std::string& v = str(k);
        return v == "1" || v == "true" || v == "True" || v == "yes" || v == "on";
    }
    std::vector<double> vec(const std::string& k) const {
        std::string s = str(k);
        for (char& c : s)
            if (c == ',' || c == '[' || c == ']') c = ' ';
        std::istringstream in(s);
        std::vector<double> out;
        double x;
        while (in >> x) out.push_back(x);
        return out;
    }
    void set(const std::string& k, const std::string& v) {
        if (!has(k)) throw std::runtime_error("unknown parameter '" + k + "' (--help lists them)");
        p[k].value = v;
    }
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n\"'"), b = s.find_last_not_of(" \t\r\n\"'");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

params define_parameters() {
    params p;
    // fk_mc<L>::define_parameters (include/fk_mc/fk_mc.hxx:176-207)
    p.define("beta", "10", "Inverse temperature");
    p.define("U", "1", "FK U");
    p.define("mu_c", "0.5", "Chemical potential of c electrons");
    p.define("mu_f", "0.5", "Chemical potential of f electrons");
    p.define("mc_flip", "0", "Make flip moves");
    p.define("mc_add_remove", "1", "Make add/remove moves");
    p.define("mc_reshuffle", "0", "Make reshuffle moves");
    p.define("cheb_moves", "false", "Allow moves using Chebyshev sampling");
    p.define("cheb_prefactor", "2.2", "Prefactor for number of Chebyshev polynomials = #ln(Volume)");
    p.define("measure_history", "true", "Measure the history");
    p.define("Nf_start", "5", "Starting number of f-electrons");
    p.define("seed", "", "Seed for random number generator (default: std::random_device)");
    p.define("measure_ipr", "false", "Measure inverse participation ratio");
    p.define("measure_eigenfunctions", "false", "Measure eigenfunctions");
    p.define("cond_offset", "0.05", "dos offset from the real axis");
    p.define("measure_stiffness", "false", "Measure stiffness/conductivity");
    // mc_metropolis::define_parameters (src/mc_metropolis.cpp:11-19)
    p.define("nsweeps", "1024", "Total number of sweeps (1 sweep = #sweep_len moves + 1 measurement)");
    p.define("sweep_len", "16", "Number of moves between subsequent measurements");
    p.define("ntherm_sweeps", "1", "How many sweeps to do before start measuring");
    p.define("show_output", "true", "Show run progress of mc");
    // cmdline_params (prog/fk_mc_exec.cpp:182-219)
    p.define("L", "4", "System linear size");
    p.define("t", "1", "Hopping");
    p.define("tp", "1", "Triangular lattice : NNN Hopping");
    p.define("W", "", "1d : f-f interaction");
    p.define("output", "output.h5", "HDF5 file to write data to, or a directory for .npy files (no .h5 suffix)");
    p.define("plaintext", "false", "also write the series as text (.dat)");
    p.define("dos_width", "6", "Width of DOS");
    p.define("dos_npts", "240", "Number of points for DOS sampling");
    p.define("dos_offset", "0.05", "DOS offset from real axis");
    p.define("cond_npoints", "150", "number of points to sample conductivity");
    p.define("save_eigenfunctions", "false", "Store eigenfunctions?");
    p.define("maxtime", "720", "max evaluation time (defined, not used, as in the reference)");
    p.define("h5lattice", "kwant_lattice.h5", "H5 for lattice (graph lattices: not built)");
    p.define("exit", "false", "dry_run: print the parameters and exit");
    p.define("max_time", "1209600", "Maximum running time in seconds");
    // batched program
    p.define("lattice", "cubic2d", "cubic1d | cubic2d | cubic3d | triangular");
    p.define("chains", "1", "Number of chains (chain r = MPI rank r of the reference, seeded seed + r)");
    p.define("mc_hop", "0", "Make nearest-neighbour exchange moves (no reference counterpart)");
    p.define("update", "dense", "dense | rank1 (rank-one spectral updates, N <= 1024)");
    p.define("refresh", "256", "rank1: full re-decomposition every this many moves");
    p.define("precision", "fp64", "fp64 | mixed");
    p.define("device", "0", "HIP device");
    return p;
}

void read_ini(params& p, const std::string& file) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot read " + file);
    std::string line;
    while (std::getline(in, line)) {
        line = trim(line.substr(0, line.find('#')));
        if (line.empty() || line[0] == '[') continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        p.set(trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
    }
}

// .npy (format 1.0): little-endian float64, C order
void write_npy(const std::string& path, const double* data, const std::vector<size_t>& shape) {
    std::string dims;
    size_t n = 1;
    for (size_t s : shape) {
        dims += std::to_string(s) + ", ";
        n *= s;
    }
    if (shape.size() > 1) dims = dims.substr(0, dims.size() - 2);
    std::string hdr = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + dims + "), }";
    while ((10 + hdr.size() + 1) % 64) hdr += ' ';
    hdr += '\n';
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write " + path);
    const char magic[] = "\x93NUMPY\x01\x00";
    out.write(magic, 8);
    const uint16_t hl = (uint16_t)hdr.size();
    out.write(reinterpret_cast<const char*>(&hl), 2);
    out.write(hdr.data(), hdr.size());
    out.write(reinterpret_cast<const char*>(data), sizeof(double) * n);
}

void write_txt(const std::string& path, const double* data, size_t rows, size_t cols) {
    std::ofstream out(path);
    out.precision(17);
    for (size_t r = 0; r < rows; ++r) {
        for (size_t c = 0; c < cols; ++c) out << (c ? " " : "") << data[r * cols + c];
        out << "\n";
    }
}

void mkdirs(const std::string& path) {
    std::string cur;
    std::istringstream in(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(in, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) throw std::runtime_error("cannot create " + cur);
    }
}

// Where the run's datasets go (prog/data_save.hxx:32-151): with `output` ending in .h5 one
// HDF5 file, dataset /<path> for each path (through libfkmc_h5.so, loaded here only then),
// as the reference writes; otherwise a directory holding <path>.npy.  With one chain the HDF5
// datasets of mc_data/ drop the chain axis, so that they have the reference's shapes.  Text
// files (parameters.txt in directory mode, stats.txt, the plaintext .dat files) go to the
// directory, or for an HDF5 file to the directory that holds it, as the reference's savetxt
// writes next to its output.h5 in the working directory.
class archive {
public:
    archive(const std::string& output, int chains) : chains_(chains) {
        const bool h5 = output.size() > 3 && output.substr(output.size() - 3) == ".h5";
        if (!h5) {
            dir_ = output;
            mkdirs(dir_);
            return;
        }
        const size_t slash = output.find_last_of('/');
        dir_ = slash == std::string::npos ? "." : output.substr(0, slash);
        if (dir_.empty()) dir_ = "/";
        mkdirs(dir_);
        char exe[4096];
        const ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
        std::string lib = "libfkmc_h5.so";
        if (n > 0) {
            exe[n] = 0;
            std::string d(exe);
            d = d.substr(0, d.find_last_of('/'));
            lib = d + "/../lib/libfkmc_h5.so";
        }
        so_ = dlopen(lib.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!so_)
            throw std::runtime_error("HDF5 output (`output` ending in .h5) needs " + lib + ", built where an HDF5 C " +
                                     "library is installed (" + dlerror() + "); give `output` a directory name for .npy output");
        create_ = (void* (*)(const char*))dlsym(so_, "fkh5_create");
        f64_ = (int (*)(void*, const char*, const double*, int, const uint64_t*))dlsym(so_, "fkh5_write_f64");
        i64_ = (int (*)(void*, const char*, const int64_t*, int, const uint64_t*))dlsym(so_, "fkh5_write_i64");
        str_ = (int (*)(void*, const char*, const char*))dlsym(so_, "fkh5_write_str");
        close_ = (int (*)(void*))dlsym(so_, "fkh5_close");
        err_ = (const char* (*)())dlsym(so_, "fkh5_error");
        if (!create_ || !f64_ || !i64_ || !str_ || !close_ || !err_) throw std::runtime_error(lib + ": missing symbols");
        h_ = create_(output.c_str());
        if (!h_) throw std::runtime_error(err_());
    }
    ~archive() {
        if (h_) close_(h_);
    }
    bool hdf5() const { return h_ != nullptr; }
    // dataset `path` (groups separated by '/'); chain_axis: shape[0] is the chain
    void put(const std::string& path, const double* data, std::vector<size_t> shape, bool chain_axis = false) {
        if (!h_) {
            const size_t slash = path.find_last_of('/');
            if (slash != std::string::npos) mkdirs(dir_ + "/" + path.substr(0, slash));
            write_npy(dir_ + "/" + path + ".npy", data, shape);
            return;
        }
        if (chain_axis && chains_ == 1 && shape.size() > 1) shape.erase(shape.begin());
        std::vector<uint64_t> d(shape.begin(), shape.end());
        if (f64_(h_, path.c_str(), data, (int)d.size(), d.data()) != 0) throw std::runtime_error(err_());
    }
    // /parameters/<name> (the reference's alps::hdf5::save of its parameters): integers as
    // int64, other numbers as float64, true / false as int64 1 / 0, anything else a string;
    // in directory mode `name = value` lines of parameters.txt
    void put_parameters(const std::vector<std::pair<std::string, std::string>>& kv) {
        if (!h_) {
            std::ofstream pf(dir_ + "/parameters.txt");
            for (auto& e : kv) pf << e.first << " = " << e.second << "\n";
            return;
        }
        for (auto& e : kv) {
            const std::string path = "parameters/" + e.first, &v = e.second;
            int r;
            char* end = nullptr;
            errno = 0;
            const long long iv = std::strtoll(v.c_str(), &end, 10);
            if (!v.empty() && *end == 0 && errno == 0) {
                const int64_t x = iv;
                r = i64_(h_, path.c_str(), &x, 0, nullptr);
            } else if (v == "true" || v == "false") {
                const int64_t x = v == "true";
                r = i64_(h_, path.c_str(), &x, 0, nullptr);
            } else {
                const double dv = std::strtod(v.c_str(), &end);
                if (!v.empty() && *end == 0)
                    r = f64_(h_, path.c_str(), &dv, 0, nullptr);
                else
                    r = str_(h_, path.c_str(), v.c_str());
            }
            if (r != 0) throw std::runtime_error(err_());
        }
    }
    // path of text file `name` (may name a group directory, kept in directory mode)
    std::string text(const std::string& name) const {
        const size_t slash = name.find_last_of('/');
        if (slash == std::string::npos) return dir_ + "/" + name;
        if (h_) return dir_ + "/" + name.substr(slash + 1);
        mkdirs(dir_ + "/" + name.substr(0, slash));
        return dir_ + "/" + name;
    }
    void close() {
        if (h_ && close_(h_) != 0) {
            h_ = nullptr;
            throw std::runtime_error(err_());
        }
        h_ = nullptr;
    }
    const std::string& dir() const { return dir_; }

private:
    int chains_;
    std::string dir_;
    void* so_ = nullptr;
    void* h_ = nullptr;
    void* (*create_)(const char*) = nullptr;
    int (*f64_)(void*, const char*, const double*, int, const uint64_t*) = nullptr;
    int (*i64_)(void*, const char*, const int64_t*, int, const uint64_t*) = nullptr;
    int (*str_)(void*, const char*, const char*) = nullptr;
    int (*close_)(void*) = nullptr;
    const char* (*err_)() = nullptr;
};

struct mean_err {
    double mean, err;
};
// mean over chains of per-chain values, error = spread over chains / sqrt(chains)
mean_err over_chains(const std::vector<double>& v) {
    const double n = (double)v.size();
    double m = 0.0;
    for (double x : v) m += x;
    m /= n;
    double var = 0.0;
    for (double x : v) var += (x - m) * (x - m);
    return {m, v.size() > 1 ? std::sqrt(var / (n - 1.0) / n) : NAN};
}

#define FK_CHECK(call, ctx)                                                                     \
    do {                                                                                        \
        int _r = (call);                                                                        \
        if (_r != FKGPU_OK) throw std::runtime_error(std::string(#call) + ": " + fkgpu_last_error(ctx)); \
    } while (0)

// std::to_string(float) with trailing zeros removed, the frequency label of save_gwr's
// dataset names (prog/data_save.hxx:568-572)
std::string wlabel(double x) {
    std::string s = std::to_string((float)x);
    return s.erase(s.find_last_not_of('0') + 1);
}

// one row of tdos_vals (prog/data_save.hxx:623-642) from the local DOS of the sites; the
// reference leaves the error columns at zero
std::vector<double> tdos_row(double w, double off, const double* ld, int N) {
    double dos = 0.0, lg = 0.0;
    for (int i = 0; i < N; ++i) dos += ld[i], lg += std::log(ld[i]);
    dos /= N;
    const double geom = std::exp(lg / N);
    return {w, off, geom, 0.0, dos, 0.0, geom / dos, 0.0};
}

void save_gwr(archive& ar, int D, int L, int N, double off, const std::vector<double>& gre,
              const std::vector<double>& gim, const std::vector<double>& grid, const std::vector<double>& ldos,
              bool plaintext) {
    const std::string dir = "stats/";
    const std::string ws = wlabel(0.0) + "_" + wlabel(off);
    std::vector<double> ld0(N);
    for (int i = 0; i < N; ++i) ld0[i] = gim[(size_t)i * N + i] / -M_PI;
    const std::vector<double> t0 = tdos_row(0.0, off, ld0.data(), N);
    ar.put(dir + "tdos" + ws, t0.data(), {8});
    ar.put(dir + "gr_full_w" + ws + "_re", gre.data(), {(size_t)N, (size_t)N});
    ar.put(dir + "gr_full_w" + ws + "_im", gim.data(), {(size_t)N, (size_t)N});
    if (plaintext) {
        write_txt(ar.text(dir + "gr_full_w" + ws + "_re.dat"), gre.data(), N, N);
        write_txt(ar.text(dir + "gr_full_w" + ws + "_im.dat"), gim.data(), N, N);
    }
    if (D >= 2) {
        // G(r_j - r_i) over the first two lattice axes (the site index runs fastest along
        // the last axis), then its 2-D discrete Fourier transform, forward, unnormalised
        std::vector<double> r2re((size_t)L * L, 0.0), r2im((size_t)L * L, 0.0);
        int per = 1;
        for (int d = 2; d < D; ++d) per *= L;
        auto pos01 = [&](int i, int& a, int& b) { a = i / (L * per); b = (i / per) % L; };
        for (int i = 0; i < N; ++i) {
            int ai, bi;
            pos01(i, ai, bi);
            for (int j = 0; j < N; ++j) {
                int aj, bj;
                pos01(j, aj, bj);
                const size_t q = (size_t)((L + aj - ai) % L) * L + (L + bj - bi) % L;
                r2re[q] += gre[(size_t)i * N + j] / N;
                r2im[q] += gim[(size_t)i * N + j] / N;
            }
        }
        std::vector<double> kre((size_t)L * L, 0.0), kim((size_t)L * L, 0.0);
        for (int k0 = 0; k0 < L; ++k0)
            for (int k1 = 0; k1 < L; ++k1) {
                double sr = 0.0, si = 0.0;
                for (int p0 = 0; p0 < L; ++p0)
                    for (int p1 = 0; p1 < L; ++p1) {
                        const double ph = -2.0 * M_PI * (double)((k0 * p0 + k1 * p1) % L) / L;
                        const double c = std::cos(ph), s = std::sin(ph);
                        const double xr = r2re[(size_t)p0 * L + p1], xi = r2im[(size_t)p0 * L + p1];
                        sr += xr * c - xi * s;
                        si += xr * s + xi * c;
                    }
                kre[(size_t)k0 * L + k1] = sr;
                kim[(size_t)k0 * L + k1] = si;
            }
        ar.put(dir + "gr_w" + ws + "_re", r2re.data(), {(size_t)L, (size_t)L});
        ar.put(dir + "gr_w" + ws + "_im", r2im.data(), {(size_t)L, (size_t)L});
        ar.put(dir + "gk_w" + ws + "_re", kre.data(), {(size_t)L, (size_t)L});
        ar.put(dir + "gk_w" + ws + "_im", kim.data(), {(size_t)L, (size_t)L});
    }
    const int npts = (int)grid.size();
    std::vector<double> tv;
    for (int q = 0; q < npts && !ldos.empty(); ++q) {
        const std::vector<double> r = tdos_row(grid[q], off, ldos.data() + (size_t)q * N, N);
        tv.insert(tv.end(), r.begin(), r.end());
    }
    if (!tv.empty()) {
        ar.put(dir + "tdos_gwr", tv.data(), {(size_t)npts, 8});
        if (plaintext) write_txt(ar.text(dir + "tdos_gwr.dat"), tv.data(), npts, 8);
    }
}

int run(params& p) {
    const std::string lat = p.str("lattice");
    int kind = FKGPU_LATTICE_HYPERCUBIC, D = 2;
    if (lat == "cubic1d") D = 1;
    else if (lat == "cubic2d") D = 2;
    else if (lat == "cubic3d") D = 3;
    else if (lat == "triangular") kind = FKGPU_LATTICE_TRIANGULAR;
    else throw std::runtime_error("unknown lattice " + lat);
    // include/fk_mc/fk_mc.hxx:59-78: the Chebyshev moves replace the ED ones; their weights
    // here are the exact dlogZ of the local estimator (fkgpu_chains_set_kpm_delta), which
    // covers moves of one or two sites
    const bool cheb = p.b("cheb_moves");
    if (cheb && p.d("mc_reshuffle") > 2.220446049250313e-16)
        throw std::runtime_error("cheb_moves: the batched program's Chebyshev weights cover moves of one or two "
                                 "sites (flip, add_remove, mc_hop), not mc_reshuffle");
    if (cheb && (p.str("precision") == "mixed" || p.str("update") == "rank1"))
        throw std::runtime_error("cheb_moves replaces the dense eigensolve: no --precision mixed / --update rank1");
    const int L = (int)p.i("L"), R = (int)p.i("chains");
    const long nsweeps = p.i("nsweeps"), sweep_len = p.i("sweep_len"), ntherm = p.i("ntherm_sweeps");
    if (R < 1 || nsweeps < 0 || sweep_len < 1 || ntherm < 0) throw std::runtime_error("invalid chains / sweeps");
    if (p.str("seed").empty()) p.set("seed", std::to_string((int)std::random_device()()));
    const long seed = p.i("seed");
    const double beta = p.d("beta"), U = p.d("U"), mu_c = p.d("mu_c"), mu_f = p.d("mu_f");
    const bool ipr = p.b("measure_ipr"), eig = p.b("measure_eigenfunctions");
    // prog/fk_mc_exec.cpp:62: the IPR and the eigenfunctions switch the histories on
    const bool history = p.b("measure_history") || ipr || eig;
    // include/fk_mc/fk_mc.hxx:115-124: Chebyshev runs measure the energy and the spectrum only
    // with measure_ipr (calc_spectrum); without the spectrum measures only occupations are kept
    const bool spectra = !cheb || ipr || eig || p.b("measure_stiffness");

    // prog/fk_mc_exec.cpp:97-117: the parameter summary, then `exit` ends a dry run before
    // anything is computed
    if (p.b("show_output") || p.b("exit")) {
        std::printf("Model parameters:\n  System size %d (%s)\n  U %g\n  t %g\n  Temperature %g\n  mu_c %g\n"
                    "  e_f %g\n  mu_f %g\n  beta %g\n", L, lat.c_str(), U, p.d("t"), 1.0 / beta, mu_c, mu_f - mu_c, mu_f,
                    beta);
        std::printf("Monte Carlo parameters:\n  sweeps %ld\n  moves per sweep %ld\n  warmup sweeps %ld\n  seed %ld\n"
                    "  flip %g, add/remove %g, reshuffle %g, hop %g\n", nsweeps, sweep_len, ntherm, seed,
                    p.d("mc_flip"), p.d("mc_add_remove"), p.d("mc_reshuffle"), p.d("mc_hop"));
        std::fflush(stdout);
    }
    if (p.b("exit")) return 0;
    // opened before the run, so that an output that cannot be written ends it at once
    archive ar(p.str("output"), R);
    fkgpu_ctx* ctx = nullptr;
    if (fkgpu_create((int)p.i("device"), &ctx) != FKGPU_OK) throw std::runtime_error("no HIP device");
    FK_CHECK(fkgpu_set_lattice(ctx, kind, D, L, p.d("t"), kind == FKGPU_LATTICE_TRIANGULAR ? p.d("tp") : 0.0), ctx);
    const int N = fkgpu_msize(ctx);
    std::vector<int32_t> f0((size_t)R * N, 0), nfs(R, (int32_t)p.i("Nf_start"));
    std::vector<double> vb(R, beta), vU(R, U), vmc(R, mu_c), vmf(R, mu_f);
    std::vector<uint32_t> seeds(R);
    for (int r = 0; r < R; ++r) seeds[r] = (uint32_t)(seed + r);  // SEED + rank (src/mc_metropolis.cpp:25)
    fkgpu_chains* ch = nullptr;
    FK_CHECK(fkgpu_chains_create(ctx, R, f0.data(), vb.data(), vU.data(), vmc.data(), vmf.data(), seeds.data(), &ch),
             ctx);
    // fk_mc::initialize: randomize_f from the Metropolis generator, then the moves in the
    // reference's order (flip, add_remove, reshuffle) with their weights
    FK_CHECK(fkgpu_chains_randomize_f(ch, nfs.data()), ctx);
    std::vector<int32_t> kinds;
    std::vector<double> w;
    const std::pair<const char*, int> mv[] = {{"mc_flip", FKGPU_MOVE_FLIP},
                                               {"mc_add_remove", FKGPU_MOVE_ADDREMOVE},
                                               {"mc_reshuffle", FKGPU_MOVE_RANDOMIZE},
                                               {"mc_hop", FKGPU_MOVE_HOP}};
    for (auto& m : mv)
        if (p.d(m.first) > 2.220446049250313e-16) {
            kinds.push_back(m.second);
            w.push_back(p.d(m.first));
        }
    if (kinds.empty()) throw std::runtime_error("No registered moves");
    FK_CHECK(fkgpu_chains_set_moves(ch, (int)kinds.size(), kinds.data(), w.data()), ctx);
    if (D == 1 && !p.str("W").empty()) {
        std::vector<double> W = p.vec("W");
        FK_CHECK(fkgpu_chains_set_ff(ch, (int)W.size(), W.data()), ctx);
    }
    if (p.str("precision") == "mixed") FK_CHECK(fkgpu_chains_set_precision(ch, FKGPU_PRECISION_MIXED), ctx);
    if (cheb) {
        // chebyshev_eval(cheb_size, grid) of fk_mc.hxx:62-65 sets the moment count of the chains'
        // stochastic logZ estimate (8 probes; no kernel damping, as calc_chebyshev); the
        // proposals' weights come from the local estimator, which needs neither
        int cheb_size = (int)(std::log((double)N) * p.d("cheb_prefactor"));
        cheb_size += cheb_size % 2;
        FK_CHECK(fkgpu_chains_set_kpm(ch, 8, std::max(2, std::min(512, cheb_size)), (uint64_t)seed, 0), ctx);
    }
    if (p.str("update") == "rank1") FK_CHECK(fkgpu_chains_set_update(ch, FKGPU_UPDATE_RANK1, (int)p.i("refresh")), ctx);
    // conductivity grid of prog/fk_mc_exec.cpp:127-139: 2 cond_npoints + 1 frequencies, log-spaced
    // from wmax e^-15 to wmax = max(8, 2U) on either side of 0
    const bool stiff = p.b("measure_stiffness");
    std::vector<double> wgrid;
    if (stiff) {
        const int nw = (int)p.i("cond_npoints");
        const double wmax = std::max(8.0, 2.0 * U);
        std::vector<double> w2(nw);
        for (int i = 0; i < nw; ++i) w2[i] = wmax * std::exp(nw > 1 ? -15.0 + 15.0 * i / (nw - 1.0) : 0.0);
        for (int i = nw - 1; i >= 0; --i) wgrid.push_back(-w2[i]);
        wgrid.push_back(0.0);
        for (int i = 0; i < nw; ++i) wgrid.push_back(w2[i]);
        FK_CHECK(fkgpu_chains_measure_conductivity(ch, (int)wgrid.size(), wgrid.data(), p.d("cond_offset")), ctx);
    }
    int flags = (history ? FKGPU_MEASURE_SPECTRUM_HISTORY | FKGPU_MEASURE_FOCC_HISTORY : 0) | (ipr ? FKGPU_MEASURE_IPR : 0) |
                (stiff ? FKGPU_MEASURE_STIFFNESS : 0) | (eig ? FKGPU_MEASURE_EIGENFUNCTIONS : 0);
    if (!spectra) flags = (history ? FKGPU_MEASURE_FOCC_HISTORY : 0) | FKGPU_MEASURE_NO_SPECTRUM;
    FK_CHECK(fkgpu_chains_measure_enable(ch, (int)nsweeps, flags), ctx);

    const bool show = p.b("show_output");
    if (show)
        std::printf("fk_mc_gpu: %s L=%d N=%d, %d chains, seeds %ld..%ld, %ld + %ld sweeps of %ld moves\n", lat.c_str(),
                    L, N, R, seed, seed + R - 1, ntherm, nsweeps, sweep_len);
    auto t0 = std::chrono::steady_clock::now();
    if (ntherm > 0) FK_CHECK(fkgpu_sweep_graph(ch, (int)(ntherm * sweep_len)), ctx);
    // measured sweeps in blocks, with a progress line per block; max_time ends them after the
    // block in which it ran out (the reference's alps::stop_callback, prog/fk_mc_exec.cpp:151)
    const long block = std::max(1L, nsweeps / 10);
    const double max_time = p.d("max_time");
    long done = 0;
    while (done < nsweeps) {
        const long n = std::min(block, nsweeps - done);
        FK_CHECK(fkgpu_sweep_measure(ch, (int)n, (int)sweep_len), ctx);
        done += n;
        if (show) std::printf("  %ld%%\n", 100 * done / std::max(1L, nsweeps)), std::fflush(stdout);
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (done < nsweeps && el >= max_time) {
            if (show) std::printf("  max_time (%g s) reached after %ld of %ld sweeps\n", max_time, done, nsweeps);
            break;
        }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    // results (of the sweeps done)
    int64_t count = 0;
    const size_t S = (size_t)done, RN = (size_t)R * N;
    std::vector<double> e(R * S), d2(R * S), ce(R * S), em(R), d2m(R), spec(RN), focc(RN);
    FK_CHECK(fkgpu_chains_energies(ch, &count, e.data(), d2.data(), ce.data()), ctx);
    FK_CHECK(fkgpu_chains_means(ch, &count, em.data(), d2m.data(), spec.data(), focc.data()), ctx);
    const bool shist = history && spectra;
    std::vector<double> sh(shist ? R * S * N : 0), ih(ipr ? R * S * N : 0);
    std::vector<int32_t> fh(history ? R * S * N : 0);
    if (history || ipr)
        FK_CHECK(fkgpu_chains_histories(ch, &count, shist ? sh.data() : nullptr, history ? fh.data() : nullptr,
                                        ipr ? ih.data() : nullptr),
                 ctx);
    const size_t NW = wgrid.size();
    std::vector<double> stf(stiff ? R * S : 0), cnd(stiff ? R * S * NW : 0), stfm(stiff ? R : 0);
    if (stiff) FK_CHECK(fkgpu_chains_stiffness(ch, &count, stf.data(), cnd.data(), stfm.data(), nullptr), ctx);
    std::vector<int64_t> nacc(R);
    FK_CHECK(fkgpu_chains_state(ch, nullptr, nullptr, nullptr, nullptr, nacc.data(), nullptr), ctx);
    // DOS of the recorded spectra (or of the mean spectra without histories)
    const int npts = (int)p.i("dos_npts");
    const double wd = p.d("dos_width");
    std::vector<double> grid(npts), dos(npts, 0.0);
    for (int i = 0; i < npts; ++i) grid[i] = -wd + 2.0 * wd * i / (1.0 * npts);
    // per-record DOS (and IPR(w) with measure_ipr) on the grid and at w = 0 (last column), for
    // the error bars of save_glocal and save_ipr (prog/data_save.hxx:283-328,486-532)
    const int nrec = shist ? (int)std::min<int64_t>(count, (int64_t)S) * R : 0;
    std::vector<double> grid0(grid);
    grid0.push_back(0.0);
    std::vector<double> drec, irec;
    if (spectra) {
        const std::vector<double>& src = nrec > 0 ? sh : spec;
        const int ns = (int)(src.size() / N);
        std::vector<double> d((size_t)ns * (npts + 1));
        if (ns > 0) {
            FK_CHECK(fkgpu_dos(ctx, ns, src.data(), npts + 1, grid0.data(), p.d("dos_offset"), d.data()), ctx);
            for (int s = 0; s < ns; ++s)
                for (int i = 0; i < npts; ++i) dos[i] += d[(size_t)s * (npts + 1) + i] / ns;
        }
        if (nrec > 0) drec.swap(d);
        if (nrec > 0 && ipr) {
            irec.resize((size_t)nrec * (npts + 1));
            FK_CHECK(fkgpu_ipr_w(ctx, nrec, sh.data(), ih.data(), npts + 1, grid0.data(), p.d("dos_offset"),
                                 irec.data()),
                     ctx);
        }
    }
    // data_saver::save_gwr (prog/data_save.hxx:72-79,535-708) on the recorded eigenvectors of
    // every chain: the full G at w = 0 with the DOS and typical DOS there, and the DOS and
    // typical DOS on the DOS grid
    const double doff = p.d("dos_offset");
    std::vector<double> evh(eig && p.b("save_eigenfunctions") ? R * S * N * N : 0);
    if (!evh.empty()) FK_CHECK(fkgpu_chains_eigenfunctions(ch, &count, evh.data()), ctx);
    const bool gwr = eig && count > 0 && S > 0;
    std::vector<double> g0re(gwr ? (size_t)N * N : 0), g0im(g0re.size()), ldos(gwr ? (size_t)npts * N : 0);
    if (gwr) {
        FK_CHECK(fkgpu_chains_green(ch, 0.0, doff, g0re.data(), g0im.data()), ctx);
        if (npts > 0) FK_CHECK(fkgpu_chains_local_dos(ch, npts, grid.data(), doff, ldos.data()), ctx);
    }
    fkgpu_chains_destroy(ch);
    fkgpu_destroy(ctx);

    const std::string md = "mc_data/";
    if (spectra) {
        ar.put(md + "energies", e.data(), {(size_t)R, S}, true);
        ar.put(md + "d2energies", d2.data(), {(size_t)R, S}, true);
        ar.put(md + "c_energies", ce.data(), {(size_t)R, S}, true);
        ar.put(md + "spectrum", spec.data(), {(size_t)R, (size_t)N}, true);
    }
    if (history) {
        if (shist) ar.put(md + "spectrum_history", sh.data(), {(size_t)R, S, (size_t)N}, true);
        std::vector<double> fd(fh.begin(), fh.end());
        ar.put(md + "focc_history", fd.data(), {(size_t)R, S, (size_t)N}, true);
    }
    if (ipr) ar.put(md + "ipr_history", ih.data(), {(size_t)R, S, (size_t)N}, true);
    if (!evh.empty()) ar.put(md + "eig_history", evh.data(), {(size_t)R, S, (size_t)N, (size_t)N}, true);
    if (gwr) save_gwr(ar, D, L, N, doff, g0re, g0im, grid, ldos, p.b("plaintext"));
    if (stiff) {
        ar.put(md + "stiffness", stf.data(), {(size_t)R, S}, true);
        ar.put(md + "cond_history", cnd.data(), {(size_t)R, S, NW}, true);
        ar.put("wgrid_conductivity", wgrid.data(), {NW});
    }
    std::vector<double> dg(2 * npts);
    for (int i = 0; i < npts; ++i) dg[2 * i] = grid[i], dg[2 * i + 1] = dos[i];
    if (spectra) ar.put("dos", dg.data(), {(size_t)npts, 2});
    if (p.b("plaintext") && spectra) {
        write_txt(ar.text("energies.dat"), e.data(), R, S);
        write_txt(ar.text("dos.dat"), dg.data(), npts, 2);
    }
    {
        std::vector<std::pair<std::string, std::string>> kv;
        for (auto& k : p.order) kv.emplace_back(k, p.str(k));
        ar.put_parameters(kv);
    }
    // statistics: per-chain means, errors from the spread over the chains
    std::vector<double> cv(R), eN(R), accr(R);
    for (int r = 0; r < R; ++r) {
        double m = 0, m2 = 0, md2 = 0;
        for (size_t s = 0; s < S; ++s) {
            const double x = e[r * S + s];
            m += x;
            m2 += x * x;
            md2 += d2[r * S + s];
        }
        m /= std::max<size_t>(S, 1);
        m2 /= std::max<size_t>(S, 1);
        md2 /= std::max<size_t>(S, 1);
        cv[r] = beta * beta * (m2 - md2 - m * m) / N;  // prog/data_save.hxx:185-190
        eN[r] = em[r] / N;
        accr[r] = (double)nacc[r] / (double)((ntherm + done) * sweep_len);
    }
    std::vector<double> nfm(R, 0.0);
    for (int r = 0; r < R; ++r)
        for (int i = 0; i < N; ++i) nfm[r] += focc[(size_t)r * N + i] / N;
    std::ofstream st(ar.text("stats.txt"));
    st.precision(12);
    auto line = [&](const char* name, const std::vector<double>& v) {
        const mean_err me = over_chains(v);
        st << name << " = " << me.mean << " +- " << me.err << "\n";
        if (show) std::printf("  %-12s %.10g +- %.3g\n", name, me.mean, me.err);
    };
    st << "# means over " << R << " chains x " << count << " measurements; errors from the spread over chains\n";
    if (spectra) {
        line("energy/N", eN);
        line("cv", cv);
    }
    line("nf/N", nfm);
    line("acceptance", accr);
    if (stiff) line("stiffness", stfm);
    // the reference's error analysis (prog/data_save.hxx:158-345,486-532) on the series of all
    // chains in chain order (its merge of the ranks' histories), binned from the last
    // measurement backwards: energies and the jackknifed specific heat, dos0 / dos_err and,
    // with measure_ipr, ipr0 / ipr_err
    if (S > 0 && count > 0 && spectra) {
        namespace bn = fkmc::binning;
        const size_t nm = (size_t)R * std::min<int64_t>(count, (int64_t)S);
        const int maxb = bn::max_bin(nm);
        auto series = [&](const double* x, size_t stride, size_t n) {  // reversed
            std::vector<double> v(n);
            for (size_t q = 0; q < n; ++q) v[q] = x[(n - 1 - q) * stride];
            return v;
        };
        auto save = [&](const std::string& name, const std::vector<bn::stats>& sv) {
            const std::vector<double> cl = bn::cor_length(sv);
            std::vector<double> t;
            for (size_t i = 0; i < sv.size(); ++i) t.insert(t.end(), {sv[i].size, sv[i].mean, sv[i].disp, sv[i].sqerror, cl[i]});
            ar.put("binning/" + name, t.data(), {sv.size(), 5});
            const int b = bn::estimate_bin(sv);
            const double row[4] = {sv[b].size, sv[b].mean, sv[b].disp, sv[b].sqerror};
            ar.put("stats/" + name, row, {4});
            st << name << " = " << sv[b].mean << " +- " << sv[b].sqerror << " (" << sv[b].size << " samples, bin "
               << b << ")\n";
            return b;
        };
        const std::vector<double> E = series(e.data(), 1, nm), D2 = series(d2.data(), 1, nm),
                                  CE = series(ce.data(), 1, nm);
        save("energy", bn::accumulate(E, maxb));
        save("d2energy", bn::accumulate(D2, maxb));
        save("c_energy", bn::accumulate(CE, maxb));
        std::vector<double> E2(nm);
        for (size_t q = 0; q < nm; ++q) E2[q] = E[q] * E[q];
        const int vol = N;
        save("cv", bn::accumulate_jackknife(
                       [beta, vol](const std::vector<double>& m) { return beta * beta * (m[1] - m[2] - m[0] * m[0]) / vol; },
                       {E, E2, D2}, maxb));
        auto err_table = [&](const std::string& name, const std::vector<double>& rec) {
            const size_t W = npts + 1;
            const int b = save(name + "0", bn::accumulate(series(rec.data() + npts, W, nm), maxb));
            std::vector<double> t;
            for (int i = 0; i < npts; ++i) {
                const bn::stats x = bn::bin(series(rec.data() + i, W, nm), b);
                t.insert(t.end(), {grid[i], x.mean, x.sqerror});
            }
            ar.put("stats/" + name + "_err", t.data(), {(size_t)npts, 3});
            if (p.b("plaintext")) write_txt(ar.text("stats/" + name + "_err.dat"), t.data(), npts, 3);
        };
        if (!drec.empty() && nm * (npts + 1) == drec.size()) err_table("dos", drec);
        if (!irec.empty()) err_table("ipr", irec);
    }
    st << "seconds = " << secs << "\nmoves_per_second = " << (double)R * (ntherm + done) * sweep_len / secs
       << "\n";
    st.close();
    ar.close();
    if (show) std::printf("  %.3f s, %.4g moves/s, output in %s\n", secs,
                          (double)R * (ntherm + done) * sweep_len / secs, p.str("output").c_str());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    params p = define_parameters();
    try {
        for (int a = 1; a < argc; ++a) {
            std::string s = argv[a];
            if (s == "--help" || s == "-h") {
                std::printf("usage: fk_mc_gpu [file.ini] [--name value | --name=value ...]\n");
                for (auto& k : p.order)
                    std::printf("  --%-24s %s (default: %s)\n", k.c_str(), p.p[k].help.c_str(), p.p[k].value.c_str());
                return 0;
            }
            if (s.rfind("--", 0) == 0) {
                s = s.substr(2);
                size_t eq = s.find('=');
                if (eq != std::string::npos) p.set(s.substr(0, eq), s.substr(eq + 1));
                else if (a + 1 < argc) p.set(s, argv[++a]);
                else p.set(s, "true");
            } else if (a == 1) {
                read_ini(p, s);
            } else {
                throw std::runtime_error("unexpected argument " + s);
            }
        }
        return run(p);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "fk_mc_gpu: %s\n", ex.what());
        return 1;
    }
}
