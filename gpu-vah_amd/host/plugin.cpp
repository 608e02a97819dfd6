// plugin.cpp -- the reference's run() plugin symbol over vah_run().
//
// rhic/src/Run.cpp hands run() three parameter blocks as void pointers
// (rhic/include/HydroPlugin.h:11).  They are read through the reference's
// layouts (include/vah_plugin.h), merged into one vah_config and handed to
// vah_run, which drives the device step with the reference's cadence and
// output files.
#include <stdio.h>

#include "../../include/vah_plugin.h"

extern "C" int vah_plugin_run(void *latticeParams, void *initCondParams, void *hydroParams,
                              const char * /*rootDirectory*/, const char *outputDir) {
    if (!latticeParams || !initCondParams || !hydroParams || !outputDir) return VAH_ERR_ARG;
    const vah_ref_lattice_params *L = static_cast<const vah_ref_lattice_params *>(latticeParams);
    const vah_ref_initcond_params *I = static_cast<const vah_ref_initcond_params *>(initCondParams);
    const vah_ref_hydro_params *H = static_cast<const vah_ref_hydro_params *>(hydroParams);
    vah_config c;
    c.numLatticePointsX = L->numLatticePointsX;
    c.numLatticePointsY = L->numLatticePointsY;
    c.numLatticePointsRapidity = L->numLatticePointsRapidity;
    c.numProperTimePoints = L->numProperTimePoints;
    c.latticeSpacingX = L->latticeSpacingX;
    c.latticeSpacingY = L->latticeSpacingY;
    c.latticeSpacingRapidity = L->latticeSpacingRapidity;
    c.latticeSpacingProperTime = L->latticeSpacingProperTime;
    c.initialConditionType = I->initialConditionType;
    c.numberOfNucleonsPerNuclei = I->numberOfNucleonsPerNuclei;
    c.initialEnergyDensity = I->initialEnergyDensity;
    c.scatteringCrossSectionNN = I->scatteringCrossSectionNN;
    c.impactParameter = I->impactParameter;
    c.fractionOfBinaryCollisions = I->fractionOfBinaryCollisions;
    c.rapidityVariance = I->rapidityVariance;
    c.rapidityMean = I->rapidityMean;
    c.initialProperTimePoint = H->initialProperTimePoint;
    c.shearViscosityToEntropyDensity = H->shearViscosityToEntropyDensity;
    c.freezeoutTemperatureGeV = H->freezeoutTemperatureGeV;
    c.initializePimunuNavierStokes = H->initializePimunuNavierStokes;
    return vah_run(&c, outputDir, 0);
}

void run(void *latticeParams, void *initCondParams, void *hydroParams, const char *rootDirectory,
         const char *outputDir) {
    int rc = vah_plugin_run(latticeParams, initCondParams, hydroParams, rootDirectory, outputDir);
    if (rc != VAH_OK) fprintf(stderr, "run: hydro failed with status %d: %s\n", rc, vah_last_error());
}
