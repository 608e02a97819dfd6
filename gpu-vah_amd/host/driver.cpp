// driver.cpp -- the reference's configuration files and run() loop on top of
// the device step.
//
// vah_config_load   rhic/src/{Lattice,InitialCondition,Hydro}Parameters.cpp and
//                   Properties.cpp: defaults, then key=value overrides.
// vah_run           rhic/src/HydroPlugin.cpp:130-242: initial state, nt+1
//                   RK2 steps, centre-cell printout and .dat output every 10
//                   steps (FileIO.cpp:14-59), average time per step.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vah_host.h"

namespace {

// libconfig scalar literal classification: an integer literal (no '.', 'e',
// 'E') is an int setting, anything else numeric a float setting.  The
// reference looks properties up with config_lookup_int / config_lookup_float,
// which do not convert between the two, so a mismatch keeps the default.
enum Kind { NONE, INT, FLOAT };

struct Setting {
    std::string name;
    Kind kind;
    long long ival;
    double fval;
};

std::string strip_comments(const std::string &text) {
    std::string out;
    size_t i = 0;
    bool in_str = false;
    while (i < text.size()) {
        char c = text[i];
        if (in_str) {
            out += c;
            if (c == '\\' && i + 1 < text.size()) out += text[++i];
            else if (c == '"') in_str = false;
            ++i;
        } else if (c == '"') {
            in_str = true;
            out += c;
            ++i;
        } else if (c == '#' || (c == '/' && i + 1 < text.size() && text[i + 1] == '/')) {
            while (i < text.size() && text[i] != '\n') ++i;
        } else if (c == '/' && i + 1 < text.size() && text[i + 1] == '*') {
            i += 2;
            while (i + 1 < text.size() && !(text[i] == '*' && text[i + 1] == '/')) ++i;
            i += 2;
        } else {
            out += c;
            ++i;
        }
    }
    return out;
}

bool parse_file(const std::string &path, std::vector<Setting> &out) {
    FILE *fp = fopen(path.c_str(), "r");
    if (!fp) return false;
    std::string text;
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, k);
    fclose(fp);
    text = strip_comments(text);
    // settings are separated by newlines, ';' or ','
    size_t pos = 0;
    while (pos < text.size()) {
        size_t end = text.find_first_of(";,\n", pos);
        if (end == std::string::npos) end = text.size();
        std::string stmt = text.substr(pos, end - pos);
        pos = end + 1;
        size_t eq = stmt.find_first_of("=:");
        if (eq == std::string::npos) continue;
        auto trim = [](std::string s) {
            size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
            return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
        };
        Setting s;
        s.name = trim(stmt.substr(0, eq));
        std::string v = trim(stmt.substr(eq + 1));
        s.kind = NONE;
        if (!v.empty() && v[0] != '"') {
            char *ep = nullptr;
            bool is_float = v.find_first_of(".eE") != std::string::npos;
            if (is_float) {
                double d = strtod(v.c_str(), &ep);
                if (ep && *ep == '\0') {
                    s.kind = FLOAT;
                    s.fval = d;
                }
            } else {
                long long iv = strtoll(v.c_str(), &ep, 0);
                if (ep && (*ep == '\0' || ((*ep == 'L') && ep[1] == '\0'))) {
                    s.kind = INT;
                    s.ival = iv;
                }
            }
        }
        if (!s.name.empty()) out.push_back(s);
    }
    return true;
}

void get_int(const std::vector<Setting> &st, const char *name, int *v, int dflt) {
    *v = dflt;
    for (const Setting &s : st)
        if (s.name == name) *v = s.kind == INT ? (int)s.ival : dflt;
}

void get_double(const std::vector<Setting> &st, const char *name, double *v, double dflt) {
    *v = dflt;
    for (const Setting &s : st)
        if (s.name == name) *v = s.kind == FLOAT ? s.fval : dflt;
}

// Threads that format one output file (the box's CPU share: at most 16).
int writer_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, h));
}

// printf's "%.*f" of v appended to out: std::to_chars in fixed notation gives
// the same correctly rounded digits as printf in the C locale (sign of zero,
// nan and inf included); a value too long for the scratch span takes printf.
inline void append_fixed(std::string &out, double v, int prec) {
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::fixed, prec);
    if (r.ec == std::errc()) {
        out.append(buf, r.ptr);
        return;
    }
    char big[400];
    int m = snprintf(big, sizeof big, "%.*f", prec, v);
    out.append(big, (size_t)std::max(0, std::min(m, (int)sizeof big - 1)));
}

// The reference's output() (FileIO.cpp:14-59): one line
// "%.3f\t%.3f\t%.3f\t%.8f\n" of x, y, eta and the value per interior cell, eta
// outermost, x fastest.  The coordinate columns repeat row after row, so each
// x column is formatted once and each row's "y\teta\t" once per row; the
// values are formatted in parallel, row blocks per thread, and the blocks
// written in order.
int write_dat(const double *var, double t, const char *dir, const char *name, const vah_config *c) {
    const int nx = c->numLatticePointsX, ny = c->numLatticePointsY, nz = c->numLatticePointsRapidity;
    std::vector<std::string> xs(nx);
    for (int i = 0; i < nx; ++i) {
        append_fixed(xs[i], (i - (nx - 1) / 2.) * c->latticeSpacingX, 3);
        xs[i].push_back('\t');
    }
    const long rows = (long)ny * nz;
    const int nt = (int)std::min<long>(writer_threads(), rows);
    std::vector<std::string> part(nt);
    std::vector<char> failed(nt, 0);  // a formatting thread that ran out of memory
    auto format = [&](int w) {
        std::string &o = part[w], yz;
        const long r0 = rows * w / nt, r1 = rows * (w + 1) / nt;
        o.reserve((size_t)(r1 - r0) * nx * 48);
        for (long r = r0; r < r1; ++r) {
            const int k = (int)(r / ny), j = (int)(r % ny);
            yz.clear();
            append_fixed(yz, (j - (ny - 1) / 2.) * c->latticeSpacingY, 3);
            yz.push_back('\t');
            append_fixed(yz, (k - (nz - 1) / 2.) * c->latticeSpacingRapidity, 3);
            yz.push_back('\t');
            const double *row = var + (size_t)k * nx * ny + (size_t)j * nx;
            for (int i = 0; i < nx; ++i) {
                o.append(xs[i]);
                o.append(yz);
                append_fixed(o, row[i], 8);
                o.push_back('\n');
            }
        }
    };
    auto work = [&](int w) {
        try {
            format(w);
        } catch (...) {
            failed[w] = 1;
        }
    };
    std::vector<std::thread> pool;
    int w = 1;
    try {
        for (; w < nt; ++w) pool.emplace_back(work, w);
    } catch (...) {  // no thread to be had: this one formats the rest
        for (int v = w; v < nt; ++v) work(v);
    }
    work(0);
    for (auto &th : pool) th.join();
    if (std::find(failed.begin(), failed.end(), 1) != failed.end()) return VAH_ERR_NOMEM;
    char fname[1024];
    snprintf(fname, sizeof fname, "%s/%s_%.3f.dat", dir, name, t);
    FILE *fp = fopen(fname, "w");
    if (!fp) return VAH_ERR_ARG;
    int rc = VAH_OK;
    for (const std::string &o : part)
        if (fwrite(o.data(), 1, o.size(), fp) != o.size()) rc = VAH_ERR_ARG;
    if (fclose(fp) != 0) rc = VAH_ERR_ARG;
    return rc;
}

void write_bin(const double *var, size_t n, double t, const char *dir, const char *name) {
    char fname[1024];
    snprintf(fname, sizeof fname, "%s/%s_%.3f.bin", dir, name, t);
    FILE *fp = fopen(fname, "wb");
    if (!fp) return;
    fwrite(var, sizeof(double), n, fp);
    fclose(fp);
}

// The files of one output step, in the order of HydroPlugin.cpp:34-108.
int write_output(const double *q, const double *e, const double *p, const double *u, size_t n, double t,
                 const char *dir, int binary, const vah_config *c) {
    static const char *names[VAH_NQ] = {"ttt",  "ttx",  "tty",  "ttn",  "pl",   "pitt", "pitx", "pity",
                                        "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"};
    int rc = VAH_OK;
    auto put = [&](const double *v, const char *name) {
        if (rc == VAH_OK) rc = write_dat(v, t, dir, name, c);
    };
    put(e, "e");
    put(p, "p");
    put(u + n, "ux");
    put(u + 2 * n, "uy");
    put(u + 3 * n, "un");
    put(u, "ut");
    put(q + VAH_PL * n, "pl");
    const int pi_order[10] = {VAH_PIXX, VAH_PIXY, VAH_PIXN, VAH_PIYY, VAH_PIYN,
                              VAH_PITT, VAH_PITX, VAH_PITY, VAH_PITN, VAH_PINN};
    for (int f : pi_order) put(q + (size_t)f * n, names[f]);
    if (binary && rc == VAH_OK) {
        write_bin(e, n, t, dir, "e");
        write_bin(p, n, t, dir, "p");
        write_bin(u, 4 * n, t, dir, "u");
        write_bin(q, VAH_NQ * n, t, dir, "q");
    }
    return rc;
}

}  // namespace

extern "C" {

int vah_write_dat(const vah_config *c, const double *var, double tau, const char *dir, const char *name) {
    if (!c || !var || !dir || !name || c->numLatticePointsX < 1 || c->numLatticePointsY < 1 ||
        c->numLatticePointsRapidity < 1)
        return VAH_ERR_ARG;
    try {
        return write_dat(var, tau, dir, name, c);
    } catch (...) {  // out of host memory for the text buffers
        return VAH_ERR_NOMEM;
    }
}

int vah_config_load(const char *dir, vah_config *c) {
    if (!c) return VAH_ERR_ARG;
    std::vector<Setting> lat, ic, hy;
    std::string d = dir ? dir : ".";
    if (!parse_file(d + "/lattice.properties", lat)) {
        fprintf(stderr, "No configuration file  %s/lattice.properties found for lattice parameters.\n", d.c_str());
        fprintf(stderr, "Using default lattice configuration parameters.\n");
    }
    if (!parse_file(d + "/ic.properties", ic)) {
        fprintf(stderr, "No configuration file  %s/ic.properties found for initial condition parameters.\n",
                d.c_str());
        fprintf(stderr, "Using default initial condition configuration parameters.\n");
    }
    if (!parse_file(d + "/hydro.properties", hy)) {
        fprintf(stderr, "No configuration file  %s/hydro.properties found for hydrodynamic parameters.\n",
                d.c_str());
        fprintf(stderr, "Using default hydrodynamic configuration parameters.\n");
    }
    get_int(lat, "numLatticePointsX", &c->numLatticePointsX, 128);
    get_int(lat, "numLatticePointsY", &c->numLatticePointsY, 128);
    get_int(lat, "numLatticePointsRapidity", &c->numLatticePointsRapidity, 64);
    get_int(lat, "numProperTimePoints", &c->numProperTimePoints, 10);
    get_double(lat, "latticeSpacingX", &c->latticeSpacingX, 0.08);
    get_double(lat, "latticeSpacingY", &c->latticeSpacingY, 0.08);
    get_double(lat, "latticeSpacingRapidity", &c->latticeSpacingRapidity, 0.3);
    get_double(lat, "latticeSpacingProperTime", &c->latticeSpacingProperTime, 0.01);
    get_int(ic, "initialConditionType", &c->initialConditionType, 2);
    get_int(ic, "numberOfNucleonsPerNuclei", &c->numberOfNucleonsPerNuclei, 208);
    get_double(ic, "initialEnergyDensity", &c->initialEnergyDensity, 1.0);
    get_double(ic, "scatteringCrossSectionNN", &c->scatteringCrossSectionNN, 62);
    get_double(ic, "impactParameter", &c->impactParameter, 7);
    get_double(ic, "fractionOfBinaryCollisions", &c->fractionOfBinaryCollisions, 0.5);
    get_double(ic, "rapidityVariance", &c->rapidityVariance, 0.5);
    get_double(ic, "rapidityMean", &c->rapidityMean, 0.5);
    get_double(hy, "initialProperTimePoint", &c->initialProperTimePoint, 0.1);
    get_double(hy, "shearViscosityToEntropyDensity", &c->shearViscosityToEntropyDensity, 0.0795775);
    get_double(hy, "freezeoutTemperatureGeV", &c->freezeoutTemperatureGeV, 0.155);
    get_int(hy, "initializePimunuNavierStokes", &c->initializePimunuNavierStokes, 1);
    return VAH_OK;
}

int vah_run(const vah_config *c, const char *outdir, int binary) {
    if (!c || !outdir) return VAH_ERR_ARG;
    const int nx = c->numLatticePointsX, ny = c->numLatticePointsY, nz = c->numLatticePointsRapidity;
    const int nt = c->numProperTimePoints;
    const size_t n = (size_t)nx * ny * nz;
    const double t0 = c->initialProperTimePoint, dt = c->latticeSpacingProperTime;
    const double hbarc = VAH_HBARC;
    const double Tf = c->freezeoutTemperatureGeV / hbarc;
    const double ef = 13.8997 * (double)powf((float)Tf, 4.0f);
    printf("Grid size = %d x %d x %d\n", nx, ny, nz);
    printf("spatial resolution = (%.3f, %.3f, %.3f)\n", c->latticeSpacingX, c->latticeSpacingY,
           c->latticeSpacingRapidity);
    printf("Grid size [fm] = %.3f x %.3f x %.3f\n", (nx - 1) / 2 * c->latticeSpacingX,
           (ny - 1) / 2 * c->latticeSpacingY, (nz - 1) / 2 * c->latticeSpacingRapidity);
    printf("freezeout temperature = %.3f [fm^-1] (eF = %.3f [fm^-4])\n", Tf, ef);
    printf("Using conformal EOS: e=3p.\n");
    printf("eta/s = %.6f\n", c->shearViscosityToEntropyDensity);

    vah_lattice L = {nx, ny, nz, nt, c->latticeSpacingX, c->latticeSpacingY, c->latticeSpacingRapidity, dt};
    vah_hydro H = {t0, c->shearViscosityToEntropyDensity, c->freezeoutTemperatureGeV,
                   c->initializePimunuNavierStokes};
    vah_ctx *ctx = nullptr;
    int rc = vah_create(&L, &H, nullptr, &ctx);
    if (rc != VAH_OK) return rc;

    std::vector<double> e(n), ut(n), ux(n), uy(n), un(n);
    printf("Setting initial conditions: type %d.\n", c->initialConditionType);
    if ((rc = vah_ic_reference(c, e.data(), ut.data(), ux.data(), uy.data(), un.data())) != VAH_OK) {
        printf("Initial condition type not defined. Exiting ...\n");
        vah_destroy(ctx);
        return rc;
    }
    printf(c->initializePimunuNavierStokes == 1 ? "Initialize \\pi^\\mu\\nu to its asymptotic Navier-Stokes value.\n"
                                                : "Initialize \\pi^\\mu\\nu to zero.\n");
    if ((rc = vah_set_primitives(ctx, e.data(), nullptr, ut.data(), ux.data(), uy.data(), un.data(), nullptr)) !=
        VAH_OK) {
        vah_destroy(ctx);
        return rc;
    }
    // centre cell (HydroPlugin.cpp:198-201) in interior coordinates
    const int ncx = nx + 4, ncy = ny + 4, ncz = nz + 4;
    const int ictr = (nx % 2 == 0) ? ncx / 2 : (ncx - 1) / 2;
    const int jctr = (ny % 2 == 0) ? ncy / 2 : (ncy - 1) / 2;
    const int kctr = (nz % 2 == 0) ? ncz / 2 : (ncz - 1) / 2;
    const size_t sctr = (size_t)(kctr - 2) * nx * ny + (size_t)(jctr - 2) * nx + (ictr - 2);

    // Output: the state of an output step is downloaded into (q, ev, pv, u),
    // then handed to a writer thread that formats and writes its files while
    // the device runs the next steps; the next output waits for it first.
    std::vector<double> q(VAH_NQ * n), ev(n), pv(n), u(4 * n);
    std::vector<double> wq(VAH_NQ * n), we(n), wp(n), wu(4 * n);
    std::thread writer;
    int wrc = VAH_OK;
    struct Joiner {
        std::thread &t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } joiner{writer};
    auto finish_writer = [&]() {
        if (writer.joinable()) writer.join();
        return wrc;
    };
    auto fail = [&](int code) {
        finish_writer();
        vah_destroy(ctx);
        return code;
    };
    double t = t0, total_ms = 0;
    int nsteps = 0;
    for (int step = 1; step <= nt + 1; ++step) {
        if ((step - 1) % 10 == 0) {
            if ((rc = vah_synchronize(ctx)) != VAH_OK ||
                (rc = vah_get_state(ctx, q.data(), ev.data(), pv.data(), u.data(), nullptr)) != VAH_OK)
                return fail(rc);
            double Tc = ev[sctr] > 0 ? (double)powf((float)(ev[sctr] / 13.8997), 0.25f) : 0.0;
            printf("n = %d:%d (t = %.3f),\t (e, p) = (%.3f, %.3f) [GeV/fm^3],\t (T = %.3f [GeV]),\t", step - 1, nt, t,
                   ev[sctr] * hbarc, pv[sctr] * hbarc, Tc * hbarc);
            if ((rc = finish_writer()) != VAH_OK) {
                fprintf(stderr, "vah_run: cannot write the output files in %s\n", outdir);
                return fail(rc);
            }
            q.swap(wq);
            ev.swap(we);
            pv.swap(wp);
            u.swap(wu);
            auto job = [&, t]() {
                try {
                    wrc = write_output(wq.data(), we.data(), wp.data(), wu.data(), n, t, outdir, binary, c);
                } catch (...) {  // out of host memory for the text buffers
                    wrc = VAH_ERR_NOMEM;
                }
            };
            try {
                writer = std::thread(job);
            } catch (...) {  // no thread to be had: write before stepping on
                job();
            }
        }
        auto a = std::chrono::steady_clock::now();
        if ((rc = vah_step(ctx, 1)) != VAH_OK || (rc = vah_synchronize(ctx)) != VAH_OK) {
            fprintf(stderr, "step %d: %s\n", step, vah_last_error());
            return fail(rc);
        }
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
        if ((step - 1) % 10 == 0) printf("(Elapsed time/step: %.3f ms)\n", ms);
        total_ms += ms;
        ++nsteps;
        t = t0 + step * dt;
    }
    if ((rc = finish_writer()) != VAH_OK) {
        fprintf(stderr, "vah_run: cannot write the output files in %s\n", outdir);
        return fail(rc);
    }
    printf("Average time/step: %.3f ms\n", total_ms / (double)nsteps);
    vah_destroy(ctx);
    return VAH_OK;
}

}  // extern "C"
