/*
 * vah_plugin.h -- the reference's hydro plugin entry point, with its own
 * signature, over the device step.
 *
 * The reference's driver (rhic/src/Run.cpp:34-36,93-96) calls
 *     void run(void *latticeParams, void *initCondParams, void *hydroParams,
 *              const char *rootDirectory, const char *outputDir);
 * declared with C++ linkage in rhic/include/HydroPlugin.h:11, passing the
 * three parameter structs its property loaders filled.  libvah.so exports
 * that symbol unchanged, so the reference's Run.cpp and loaders link against
 * it as they link against HydroPlugin.cpp, and a C-linkage twin for FFI
 * callers.  Both read the three blocks through the layouts below, which are
 * those of the reference's structs on x86-64.
 */
#ifndef VAH_PLUGIN_H_
#define VAH_PLUGIN_H_

#include "vah_host.h"

/* struct LatticeParameters, rhic/include/LatticeParameters.h:17-31
 * (the computational sizes are the loader's nx + 4 etc.; run() ignores them) */
typedef struct vah_ref_lattice_params {
    int numLatticePointsX, numLatticePointsY, numLatticePointsRapidity;
    int numComputationalLatticePointsX, numComputationalLatticePointsY, numComputationalLatticePointsRapidity;
    int numProperTimePoints;
    double latticeSpacingX, latticeSpacingY, latticeSpacingRapidity, latticeSpacingProperTime;
} vah_ref_lattice_params;

/* struct InitialConditionParameters, rhic/include/InitialConditionParameters.h:13-27 */
typedef struct vah_ref_initcond_params {
    int initialConditionType;
    int numberOfNucleonsPerNuclei;
    double initialEnergyDensity, scatteringCrossSectionNN, impactParameter, fractionOfBinaryCollisions;
    double rapidityVariance, rapidityMean;
} vah_ref_initcond_params;

/* struct HydroParameters, rhic/include/HydroParameters.h:15-21 */
typedef struct vah_ref_hydro_params {
    double initialProperTimePoint, shearViscosityToEntropyDensity, freezeoutTemperatureGeV;
    int initializePimunuNavierStokes;
} vah_ref_hydro_params;

#ifdef __cplusplus
extern "C" {
#endif

/* The C-linkage twin of run(): the same arguments and behaviour (the
 * reference IC of the given type, nt+1 steps, the centre-cell printout and
 * the .dat files every 10 steps in outputDir; rootDirectory is accepted and
 * unused, as in the reference's run()), returning a status where run()
 * returns nothing.  NULL parameter blocks or outputDir: VAH_ERR_ARG. */
int vah_plugin_run(void *latticeParams, void *initCondParams, void *hydroParams, const char *rootDirectory,
                   const char *outputDir);

#ifdef __cplusplus
}

/* rhic/include/HydroPlugin.h:11, C++ linkage as the reference declares it.
 * Replaces run() of rhic/src/HydroPlugin.cpp:130-242.  An error is printed
 * to stderr with its status and run() returns (the reference's CPU path
 * exit(-1)s instead, its GPU path runs on). */
void run(void *latticeParams, void *initCondParams, void *hydroParams, const char *rootDirectory,
         const char *outputDir);
#endif

#endif /* VAH_PLUGIN_H_ */
