// A caller of the plugin symbol in the shape of the reference's driver
// (rhic/src/Run.cpp:34-36,93-96): it declares run() exactly as
// rhic/include/HydroPlugin.h:11 does (C++ linkage, no library header),
// fills the reference's three parameter structs (LatticeParameters.h:17-31,
// InitialConditionParameters.h:13-27, HydroParameters.h:15-21, restated here
// with the reference's member names and order) and calls it.  Linking it
// against libvah.so shows that the mangled symbol resolves; running it
// writes the reference's output files.
//   plugin_caller OUTDIR NX NY NT ICTYPE
#include <stdio.h>
#include <stdlib.h>

struct LatticeParameters {
    int numLatticePointsX;
    int numLatticePointsY;
    int numLatticePointsRapidity;
    int numComputationalLatticePointsX;
    int numComputationalLatticePointsY;
    int numComputationalLatticePointsRapidity;
    int numProperTimePoints;
    double latticeSpacingX;
    double latticeSpacingY;
    double latticeSpacingRapidity;
    double latticeSpacingProperTime;
};

struct InitialConditionParameters {
    int initialConditionType;
    int numberOfNucleonsPerNuclei;
    double initialEnergyDensity;
    double scatteringCrossSectionNN;
    double impactParameter;
    double fractionOfBinaryCollisions;
    double rapidityVariance;
    double rapidityMean;
};

struct HydroParameters {
    double initialProperTimePoint;
    double shearViscosityToEntropyDensity;
    double freezeoutTemperatureGeV;
    int initializePimunuNavierStokes;
};

void run(void *latticeParams, void *initCondParams, void *hydroParams, const char *rootDirectory,
         const char *outputDir);

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s OUTDIR NX NY NT ICTYPE\n", argv[0]);
        return 2;
    }
    LatticeParameters lat;
    lat.numLatticePointsX = atoi(argv[2]);
    lat.numLatticePointsY = atoi(argv[3]);
    lat.numLatticePointsRapidity = 1;
    lat.numComputationalLatticePointsX = lat.numLatticePointsX + 4;
    lat.numComputationalLatticePointsY = lat.numLatticePointsY + 4;
    lat.numComputationalLatticePointsRapidity = lat.numLatticePointsRapidity + 4;
    lat.numProperTimePoints = atoi(argv[4]);
    lat.latticeSpacingX = 0.3;
    lat.latticeSpacingY = 0.3;
    lat.latticeSpacingRapidity = 0.1;
    lat.latticeSpacingProperTime = 0.01;
    InitialConditionParameters ic;
    ic.initialConditionType = atoi(argv[5]);
    ic.numberOfNucleonsPerNuclei = 197;
    ic.initialEnergyDensity = 15.6269;
    ic.scatteringCrossSectionNN = 42.0;
    ic.impactParameter = 7.0;
    ic.fractionOfBinaryCollisions = 0.5;
    ic.rapidityVariance = 0.5;
    ic.rapidityMean = 0.5;
    HydroParameters hydro;
    hydro.initialProperTimePoint = 0.5;
    hydro.shearViscosityToEntropyDensity = 0.2;
    hydro.freezeoutTemperatureGeV = 0.140;
    hydro.initializePimunuNavierStokes = 0;
    run(&lat, &ic, &hydro, ".", argv[1]);
    printf("Done hydro.\n");
    return 0;
}
