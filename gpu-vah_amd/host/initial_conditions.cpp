// initial_conditions.cpp -- energy-density profiles for the initial state.
//
// Two families:
//  * parametrised Glauber geometries (optical and Monte-Carlo) used by the
//    benchmark workloads, with this library's own seeded generator;
//  * the reference's own initial conditions by initialConditionType
//    (rhic/src/InitialConditions.cpp:419-455, GlauberModel.cpp,
//    MonteCarloGlauberModel.cpp), reproduced with their constants and quirks
//    for the drop-in driver.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "../../include/vah_host.h"
#include "ic_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---- Gauss-Legendre nodes on [-1, 1] -------------------------------------
struct Gauss {
    std::vector<double> x, w;
    explicit Gauss(int n) : x(n), w(n) {
        for (int i = 0; i < n; ++i) {
            double z = cos(kPi * (i + 0.75) / (n + 0.5)), pp = 0;
            for (int it = 0; it < 100; ++it) {
                double p1 = 1, p2 = 0;
                for (int j = 0; j < n; ++j) {
                    double p3 = p2;
                    p2 = p1;
                    p1 = ((2 * j + 1) * z * p2 - j * p3) / (j + 1);
                }
                pp = n * (z * p1 - p2) / (z * z - 1);
                double dz = p1 / pp;
                z -= dz;
                if (fabs(dz) < 1e-16) break;
            }
            x[i] = z;
            w[i] = 2 / ((1 - z * z) * pp * pp);
        }
    }
};

// integral of f over [a, b] with `panels` Gauss-Legendre panels
template <typename F>
double integrate(F f, double a, double b, int panels, const Gauss &g) {
    double s = 0, h = (b - a) / panels;
    for (int k = 0; k < panels; ++k) {
        double c = a + (k + 0.5) * h;
        for (size_t i = 0; i < g.x.size(); ++i) s += g.w[i] * f(c + 0.5 * h * g.x[i]);
    }
    return 0.5 * h * s;
}

struct WoodsSaxon {
    double n0, R, a;
    double rho(double r) const { return n0 / (1 + exp((r - R) / a)); }
    double zmax() const { return R + 40 * a; }
    // nuclear thickness T(r_perp) = int rho(sqrt(r_perp^2 + z^2)) dz over all z
    double thickness(double rp, const Gauss &g) const {
        return 2 * integrate([&](double z) { return rho(sqrt(rp * rp + z * z)); }, 0.0, zmax(), 8, g);
    }
};

WoodsSaxon normalised_ws(int A, double R, double a, const Gauss &g) {
    WoodsSaxon ws{1.0, R, a};
    double vol = 4 * kPi * integrate([&](double r) { return r * r * ws.rho(r); }, 0.0, ws.zmax(), 16, g);
    ws.n0 = A / vol;
    return ws;
}

// thickness on the lattice columns for a nucleus centred at (xc, 0)
std::vector<double> thickness_map(const vah_lattice &L, const WoodsSaxon &ws, double xc, const Gauss &g) {
    std::vector<double> T((size_t)L.nx * L.ny);
    for (int j = 0; j < L.ny; ++j) {
        double y = (j - (L.ny - 1) / 2.) * L.dy;
        for (int i = 0; i < L.nx; ++i) {
            double x = (i - (L.nx - 1) / 2.) * L.dx - xc;
            T[(size_t)j * L.nx + i] = ws.thickness(sqrt(x * x + y * y), g);
        }
    }
    return T;
}

void extrude(const vah_lattice &L, const std::vector<double> &eT, double eta_flat, double eta_sigma, double floor_,
             double *e) {
    for (int k = 0; k < L.nz; ++k) {
        double eta = (k - (L.nz - 1) / 2.) * L.dz;
        double fl = L.nz > 1 ? vah_ic_eta_profile(eta, eta_flat, eta_sigma) : 1.0;
        for (size_t m = 0; m < eT.size(); ++m) e[(size_t)k * eT.size() + m] = eT[m] * fl + floor_;
    }
}

void scale_to_peak(std::vector<double> &eT, double peak) {
    double mx = *std::max_element(eT.begin(), eT.end());
    if (mx > 0)
        for (double &v : eT) v *= peak / mx;
}

// ---- xoshiro256** seeded through splitmix64 -------------------------------
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; ++i) {
            seed += 0x9e3779b97f4a7c15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * 0x1.0p-53; }
};

// A nucleon positions (transverse) from a Woods-Saxon nucleus centred at xc
void sample_nucleus(Rng &rng, int A, double R, double a, double xc, std::vector<double> &x, std::vector<double> &y) {
    const double rmax = R + 12 * a;
    // envelope of r^2 / (1 + exp((r-R)/a)) on [0, rmax]
    double M = 0;
    for (int i = 0; i <= 4000; ++i) {
        double r = rmax * i / 4000.0;
        M = std::max(M, r * r / (1 + exp((r - R) / a)));
    }
    M *= 1.01;
    x.clear();
    y.clear();
    while ((int)x.size() < A) {
        double r = rmax * rng.uniform();
        if (rng.uniform() * M < r * r / (1 + exp((r - R) / a))) {
            double ct = 2 * rng.uniform() - 1, ph = 2 * kPi * rng.uniform();
            double st = sqrt(1 - ct * ct);
            x.push_back(r * st * cos(ph) + xc);
            y.push_back(r * st * sin(ph));
        }
    }
}

double conformal_e_of_T(double T) { return 13.8997 * (double)powf((float)T, 4.0f); }

}  // namespace

namespace vah_host {

int mc_glauber_nucleons(const vah_glauber &g, std::vector<double> &xa, std::vector<double> &ya,
                        std::vector<double> &xb, std::vector<double> &yb) {
    if (g.A < 1 || !(g.R > 0) || !(g.a > 0)) return VAH_ERR_ARG;
    Rng rng(g.seed);
    sample_nucleus(rng, g.A, g.R, g.a, -g.b / 2, xa, ya);
    sample_nucleus(rng, g.A, g.R, g.a, +g.b / 2, xb, yb);
    return VAH_OK;
}

}  // namespace vah_host

extern "C" {

double vah_ic_eta_profile(double eta, double eta_flat, double eta_sigma) {
    double a = fabs(eta) - eta_flat;
    return a > 0 ? exp(-a * a / (2 * eta_sigma * eta_sigma)) : 1.0;
}

int vah_ic_optical_glauber(const vah_lattice *lat, const vah_glauber *g, double *e) {
    if (!lat || !g || !e || lat->nx < 1 || lat->ny < 1 || lat->nz < 1 || g->A < 1 || !(g->R > 0) || !(g->a > 0))
        return VAH_ERR_ARG;
    Gauss gl(64);
    WoodsSaxon ws = normalised_ws(g->A, g->R, g->a, gl);
    std::vector<double> TA = thickness_map(*lat, ws, -g->b / 2, gl);
    std::vector<double> TB = thickness_map(*lat, ws, +g->b / 2, gl);
    const double sig = 0.1 * g->sigma_nn_mb;  // fm^2
    std::vector<double> eT(TA.size());
    for (size_t m = 0; m < TA.size(); ++m) {
        double ta = TA[m], tb = TB[m];
        double ncoll = sig * ta * tb;
        double npart = ta * (1 - pow(1 - sig * tb / g->A, g->A)) + tb * (1 - pow(1 - sig * ta / g->A, g->A));
        eT[m] = (1 - g->alpha) * npart + g->alpha * ncoll;
    }
    scale_to_peak(eT, g->e_peak);
    extrude(*lat, eT, g->eta_flat, g->eta_sigma, g->e_floor, e);
    return VAH_OK;
}

int vah_ic_mc_glauber(const vah_lattice *lat, const vah_glauber *g, double *e, int *n_wounded) {
    if (!lat || !g || !e || lat->nx < 1 || lat->ny < 1 || lat->nz < 1 || g->A < 1 || !(g->R > 0) || !(g->a > 0) ||
        !(g->nucleon_width > 0))
        return VAH_ERR_ARG;
    std::vector<double> xa, ya, xb, yb;
    vah_host::mc_glauber_nucleons(*g, xa, ya, xb, yb);
    const double d2 = 0.1 * g->sigma_nn_mb / kPi;
    std::vector<char> wa(g->A, 0), wb(g->A, 0);
    for (int i = 0; i < g->A; ++i)
        for (int j = 0; j < g->A; ++j) {
            double dx = xa[i] - xb[j], dy = ya[i] - yb[j];
            if (dx * dx + dy * dy < d2) wa[i] = wb[j] = 1;
        }
    std::vector<double> px, py;
    for (int i = 0; i < g->A; ++i) {
        if (wa[i]) { px.push_back(xa[i]); py.push_back(ya[i]); }
        if (wb[i]) { px.push_back(xb[i]); py.push_back(yb[i]); }
    }
    if (n_wounded) *n_wounded = (int)px.size();
    std::vector<double> eT((size_t)lat->nx * lat->ny, 0.0);
    const double inv2w2 = 1 / (2 * g->nucleon_width * g->nucleon_width);
    const double cut2 = 64 * g->nucleon_width * g->nucleon_width;  // deposits beyond 8 widths < 1e-13
    for (int j = 0; j < lat->ny; ++j) {
        double y = (j - (lat->ny - 1) / 2.) * lat->dy;
        for (int i = 0; i < lat->nx; ++i) {
            double x = (i - (lat->nx - 1) / 2.) * lat->dx, s = 0;
            for (size_t n = 0; n < px.size(); ++n) {
                double r2 = (x - px[n]) * (x - px[n]) + (y - py[n]) * (y - py[n]);
                if (r2 < cut2) s += exp(-r2 * inv2w2);
            }
            eT[(size_t)j * lat->nx + i] = s;
        }
    }
    scale_to_peak(eT, g->e_peak);
    extrude(*lat, eT, g->eta_flat, g->eta_sigma, g->e_floor, e);
    return VAH_OK;
}

// ---- the reference's initial conditions -----------------------------------

int vah_ic_reference(const vah_config *cfg, double *e, double *ut, double *ux, double *uy, double *un) {
    if (!cfg || !e) return VAH_ERR_ARG;
    const int nx = cfg->numLatticePointsX, ny = cfg->numLatticePointsY, nz = cfg->numLatticePointsRapidity;
    const double dx = cfg->latticeSpacingX, dy = cfg->latticeSpacingY, dz = cfg->latticeSpacingRapidity;
    const size_t n = (size_t)nx * ny * nz, nT = (size_t)nx * ny;
    if (nx < 1 || ny < 1 || nz < 1) return VAH_ERR_ARG;
    for (size_t m = 0; m < n; ++m) {
        if (ut) ut[m] = 1.0;
        if (ux) ux[m] = 0.0;
        if (uy) uy[m] = 0.0;
        if (un) un[m] = 0.0;
    }
    // longitudinal profile with the reference's hard-coded plateau 5.9 and
    // variance 0.16 and its integer centre (nz-1)/2 (InitialConditions.cpp:196-215)
    std::vector<double> eL(nz);
    for (int k = 0; k < nz; ++k) {
        double eta = (k - (nz - 1) / 2) * dz;
        double s = fabs(eta) - 5.9 / 2;
        eL[k] = s < 0 ? 1.0 : exp(-s * s / 0.16 / 2);
    }
    const double A = cfg->numberOfNucleonsPerNuclei;
    // Woods-Saxon with n0 = 0.17, Rn = 1.12 A^1/3 - 0.86 A^-1/3, d = 0.54 (GlauberModel.cpp:19-24)
    const WoodsSaxon ws{0.17, 1.12 * pow(A, 1. / 3.) - 0.86 * pow(A, -1. / 3.), 0.54};
    switch (cfg->initialConditionType) {
    case 0: {  // constant e(T0 = 3.05) (InitialConditions.cpp:168-191)
        double ed = conformal_e_of_T(3.05);
        for (size_t m = 0; m < n; ++m) e[m] = ed;
        return VAH_OK;
    }
    case 2: {  // optical Glauber (InitialConditions.cpp:220-253, GlauberModel.cpp:71-108)
        Gauss gl(64);
        const double b = cfg->impactParameter, snn = cfg->scatteringCrossSectionNN;
        const double alpha = cfg->fractionOfBinaryCollisions;
        double T0 = ws.thickness(0.0, gl);
        // wounded-nucleon term as GlauberModel.cpp:56-66 writes it: the power
        // acts on sigma T / B itself, so the (1 - ...) factor is ~1
        auto wounded = [&](double TAplus, double TBminus) {
            double part = 1 - snn * TBminus * 0.1 / A;
            double res = TAplus * (1 - pow(1 - part, A));
            part = 1 - snn * TAplus * 0.1 / A;
            res += TBminus * (1 - pow(1 - part, A));
            return res;
        };
        double nbcNorm = 1. / (snn * T0 * T0), wnNorm = 1. / wounded(T0, T0);
        double e0 = conformal_e_of_T(3.05);
        for (int j = 0; j < ny; ++j) {
            double y = (j - (ny - 1) / 2.) * dy;
            for (int i = 0; i < nx; ++i) {
                double x = (i - (nx - 1) / 2.) * dx;
                double TAm = ws.thickness(sqrt((x - b / 2) * (x - b / 2) + y * y), gl);
                double TAp = ws.thickness(sqrt((x + b / 2) * (x + b / 2) + y * y), gl);
                double ed = nbcNorm * snn * TAp * TAm * alpha + (1 - alpha) * wnNorm * wounded(TAp, TAm);
                for (int k = 0; k < nz; ++k) e[(size_t)k * nT + (size_t)j * nx + i] = e0 * ed * eL[k] + 1.e-3;
            }
        }
        return VAH_OK;
    }
    case 3: {  // ideal Gubser flow (InitialConditions.cpp:367-407)
        const double e0 = cfg->initialEnergyDensity;
        for (int j = 0; j < ny; ++j) {
            double y = (j - (ny - 1) / 2.) * dy;
            for (int i = 0; i < nx; ++i) {
                double x = (i - (nx - 1) / 2.) * dx;
                double T = 1.9048812623618392 /
                           pow(1 + pow(1 - x * x - y * y, 2) + 2 * (1 + x * x + y * y), 0.3333333333333333);
                double r = sqrt(x * x + y * y);
                double phi = atanh(2 * r / (2 + x * x + y * y));
                // the reference divides by r = 0 at the centre; there the flow is 0
                double vx = r > 0 ? sinh(phi) * x / r : 0.0, vy = r > 0 ? sinh(phi) * y / r : 0.0;
                for (int k = 0; k < nz; ++k) {
                    size_t m = (size_t)k * nT + (size_t)j * nx + i;
                    e[m] = e0 * pow(T, 4);
                    if (ux) ux[m] = vx;
                    if (uy) uy[m] = vy;
                    if (ut) ut[m] = sqrt(1 + vx * vx + vy * vy);
                }
            }
        }
        return VAH_OK;
    }
    case 4: {  // MC Glauber with the C library's rand() (MonteCarloGlauberModel.cpp:17-108)
        const int NA = cfg->numberOfNucleonsPerNuclei;
        const double b = cfg->impactParameter, snn = cfg->scatteringCrossSectionNN;
        srand(1328398221);
        auto sample = [&](std::vector<double> &x, std::vector<double> &y) {
            const double M = 1.01 * 4.5310551374155095, rmax = 20;
            int m = 0;
            while (m < NA) {
                double r = rmax * ((double)rand()) / ((double)RAND_MAX);
                double u = ((double)rand()) / ((double)RAND_MAX);
                if (u < r * r * ws.rho(r) / M) {
                    double v = 2 * (((double)rand()) / ((double)RAND_MAX) - 0.5);
                    double f = 2 * M_PI * ((double)rand()) / ((double)RAND_MAX);
                    x[m] = r * sqrt(1 - v * v) * cos(f);
                    y[m] = r * sqrt(1 - v * v) * sin(f);
                    m++;
                }
            }
        };
        std::vector<double> x1(NA), y1(NA), x2(NA), y2(NA), xp, yp;
        sample(x1, y1);
        sample(x2, y2);
        for (int i = 0; i < NA; ++i) {
            x1[i] -= b / 2;
            x2[i] += b / 2;
        }
        double dn = sqrt(0.1 * snn / M_PI);
        std::vector<int> l1(NA, 0), l2(NA, 0);
        for (int i = 0; i < NA; ++i)
            for (int j = 0; j < NA; ++j) {
                double dist = sqrt(pow(x1[i] - x2[j], 2) + pow(y1[i] - y2[j], 2));
                if (dist < dn) {
                    l1[i] += 1;
                    l2[j] += 1;
                }
            }
        for (int i = 0; i < NA; ++i) {
            if (l1[i] > 0) { xp.push_back(x1[i]); yp.push_back(y1[i]); }
            if (l2[i] > 0) { xp.push_back(x2[i]); yp.push_back(y2[i]); }
        }
        const double SIG0 = 0.46;
        double e0 = conformal_e_of_T(2.03);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double s = 0;
                for (size_t p = 0; p < xp.size(); ++p) {
                    double x = (i - ((double)nx - 1.) / 2.) * dx - xp[p];
                    double y = (j - ((double)ny - 1.) / 2.) * dy - yp[p];
                    s += exp(-x * x / 2 / SIG0 / SIG0 - y * y / 2 / SIG0 / SIG0);
                }
                for (int k = 0; k < nz; ++k) e[(size_t)k * nT + (size_t)j * nx + i] = e0 * s * eL[k] + 1.e-3;
            }
        return VAH_OK;
    }
    default:
        return VAH_ERR_ARG;
    }
}

}  // extern "C"
