/*
 * vah_host.h -- host-side C ABI around the device step: equation-of-state
 * tables, initial conditions, the reference's configuration files and its
 * run() driver.  Implemented in gpu-vah_amd/host/ and exported by libvah.so.
 * Reference paths are relative to the reference repository root.
 */
#ifndef VAH_HOST_H_
#define VAH_HOST_H_

#include "vah.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VAH_HBARC 0.197326938 /* GeV fm (rhic/src/HydroPlugin.cpp:155) */

/* ---- Equation of state ------------------------------------------------- */

/* Lattice-QCD-like crossover EOS (not in the reference; used for the
 * 201x201x101 workload): p(T) = T^4 (pi^2/90) [g_h + (g_q - g_h) S(T)],
 * S = (1 + tanh((T - Tc)/dT))/2, e = T dp/dT - p, with g_q = 42.25 (the
 * reference's conformal factor 13.8997 = pi^2 g_q / 30), g_h = 5,
 * Tc = 155 MeV, dT = 25 MeV.  Fills n nodes uniform in ln e over
 * [e_min, e_max] (fm^-4): p[i], T[i] at ln e_i = *lne0 + i * *h. */
int vah_eos_lqcd_table(int n, double e_min, double e_max, double *p, double *T, double *lne0, double *h);

/* The crossover EOS evaluated directly at temperature T (fm^-1): e and p.
 * Not in the reference (its table generator's closed form). */
void vah_eos_lqcd_eval(double T, double *e, double *p);

/* ---- Initial conditions ------------------------------------------------ */

/* Nuclear-collision geometry for the Glauber initial conditions
 * (rhic/src/GlauberModel.cpp:19-108, MonteCarloGlauberModel.cpp:17-108,
 * rhic/src/InitialConditions.cpp:196-292). */
typedef struct vah_glauber {
    int A;                    /* nucleons per nucleus (numberOfNucleonsPerNuclei) */
    double R, a;              /* Woods-Saxon radius and diffuseness [fm] */
    double sigma_nn_mb;       /* inelastic NN cross-section [mb] (scatteringCrossSectionNN) */
    double b;                 /* impact parameter [fm] (impactParameter) */
    double alpha;             /* binary-collision fraction (fractionOfBinaryCollisions):
                                 e_T ~ (1 - alpha) N_part + alpha N_coll */
    double e_peak;            /* transverse profile scaled so its lattice maximum is e_peak [fm^-4] */
    double nucleon_width;     /* MC: Gaussian width of a wounded nucleon's deposit [fm] */
    unsigned long long seed;  /* MC: RNG seed */
    double eta_flat;          /* longitudinal plateau half-width in eta_s */
    double eta_sigma;         /* width of its Gaussian tails */
    double e_floor;           /* added to every cell [fm^-4] (the reference adds 1e-3) */
} vah_glauber;

/* Optical (continuous) Glauber: interior e array (x fastest).  Replaces
 * rhic/src/GlauberModel.cpp:19-108, with an own quadrature of the thickness
 * integral in place of GSL's QAGS (GlauberModel.cpp:36-50). */
int vah_ic_optical_glauber(const vah_lattice *lat, const vah_glauber *g, double *e);

/* Monte-Carlo Glauber (rhic/src/MonteCarloGlauberModel.cpp:17-108, which
 * draws from the C library's rand()) with this library's seeded generator
 * (xoshiro256**):
 * nucleons sampled from Woods-Saxon, wounded by the black-disk criterion
 * d < sqrt(sigma/pi), each depositing a normalised-to-peak Gaussian.
 * n_wounded may be NULL. */
int vah_ic_mc_glauber(const vah_lattice *lat, const vah_glauber *g, double *e, int *n_wounded);

/* Longitudinal factor exp(-(|eta|-eta_flat)^2/(2 sigma^2)) beyond the plateau
 * (InitialConditions.cpp:196-215 with its constants as parameters). */
double vah_ic_eta_profile(double eta, double eta_flat, double eta_sigma);

/* ---- The reference's configuration and driver -------------------------- */

/* rhic/include/LatticeParameters.h, InitialConditionParameters.h,
 * HydroParameters.h: the three property sets with the loaders' defaults. */
typedef struct vah_config {
    /* lattice.properties (rhic/src/LatticeParameters.cpp:30-39) */
    int numLatticePointsX, numLatticePointsY, numLatticePointsRapidity, numProperTimePoints;
    double latticeSpacingX, latticeSpacingY, latticeSpacingRapidity, latticeSpacingProperTime;
    /* ic.properties (rhic/src/InitialConditionParameters.cpp:33-41) */
    int initialConditionType, numberOfNucleonsPerNuclei;
    double initialEnergyDensity, scatteringCrossSectionNN, impactParameter, fractionOfBinaryCollisions;
    double rapidityVariance, rapidityMean;
    /* hydro.properties (rhic/src/HydroParameters.cpp:25-29) */
    double initialProperTimePoint, shearViscosityToEntropyDensity, freezeoutTemperatureGeV;
    int initializePimunuNavierStokes;
} vah_config;

/* Fills the defaults, then overrides from <dir>/lattice.properties,
 * ic.properties and hydro.properties where present (libconfig "key=value"
 * syntax: '#' and '//' comments, optional ';', quoted strings ignored).
 * Missing files keep the defaults, as the reference's loaders do. */
int vah_config_load(const char *dir, vah_config *cfg);

/* Reference initial condition by initialConditionType
 * (InitialConditions.cpp:419-455): 0 constant e(T0 = 3.05 fm^-1),
 * 2 optical Glauber, 3 ideal Gubser, 4 MC-Glauber with srand(1328398221) and
 * the host C library's rand(); u^mu = (1,0,0,0) except for 3; pi = 0.
 * Interior arrays; ut/ux/uy/un may be NULL for types with no flow. */
int vah_ic_reference(const vah_config *cfg, double *e, double *ut, double *ux, double *uy, double *un);

/* One interior field (x fastest, then y, then eta) as <dir>/<name>_<tau %.3f>.dat
 * in the reference's text format (output(), rhic/src/FileIO.cpp:14-59): a line
 * "%.3f\t%.3f\t%.3f\t%.8f\n" of x, y, eta and the value per cell, eta outermost.
 * Byte for byte printf's output; formatted on up to 16 host threads.
 * VAH_ERR_ARG for a bad argument or a file that cannot be written. */
int vah_write_dat(const vah_config *cfg, const double *var, double tau, const char *dir, const char *name);

/* run() of rhic/src/HydroPlugin.cpp:130-242 on the device: initial state from
 * the reference IC, nt+1 steps, every 10 steps the centre-cell printout and
 * the <name>_<tau %.3f>.dat files (FileIO.cpp:14-59) in output_dir, written
 * by a host thread while the device runs the next steps.
 * binary_dumps != 0 also writes full-precision <name>_<tau>.bin files. */
int vah_run(const vah_config *cfg, const char *output_dir, int binary_dumps);

#ifdef __cplusplus
}
#endif

#endif /* VAH_HOST_H_ */
