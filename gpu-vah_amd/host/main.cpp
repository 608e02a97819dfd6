// main.cpp -- the gpu-vah executable: the reference's command line
// (rhic/src/Run.cpp:38-102, CommandLineArguments.cpp:17-107) over vah_run().
//   gpu-vah --config rhic-conf -o output -h
// -c/--config DIR, -o/--output DIR, -h/--hydro (run hydro), -t/--test
// (accepted; the reference's test runner is disabled too).
#include <argp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/vah_host.h"

namespace {

struct Cli {
    bool runTest, runHydro, binary;
    const char *configDirectory, *outputDirectory;
};

argp_option options[] = {{"test", 't', "RUN_TEST", OPTION_ARG_OPTIONAL, "Run software tests", 0},
                         {"hydro", 'h', "RUN_HYDRO", OPTION_ARG_OPTIONAL, "Run hydrodynamic simulation", 0},
                         {"output", 'o', "OUTPUT_DIRECTORY", 0, "Path to output directory", 0},
                         {"config", 'c', "CONFIG_DIRECTORY", 0, "Path to configuration directory", 0},
                         {"binary", 'b', 0, 0, "Also write full-precision binary dumps", 0},
                         {0, 0, 0, 0, 0, 0}};

error_t parse_opt(int key, char *arg, argp_state *state) {
    Cli *cli = (Cli *)state->input;
    switch (key) {
    case 't': cli->runTest = true; break;
    case 'h': cli->runHydro = true; break;
    case 'o': cli->outputDirectory = arg; break;
    case 'c': cli->configDirectory = arg; break;
    case 'b': cli->binary = true; break;
    default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

char doc[] = "Run -- A program to run a single viscous hydrodynamic simulation of a relativistic heavy ion collision";
char args_doc[] = "ARG1 ARG2";
argp parser = {options, parse_opt, args_doc, doc, 0, 0, 0};

}  // namespace

int main(int argc, char **argv) {
    Cli cli = {false, false, false, nullptr, nullptr};
    argp_parse(&parser, argc, argv, 0, 0, &cli);
    printf("configDirectory = %s\n", cli.configDirectory ? cli.configDirectory : "(null)");
    printf("outputDirectory = %s\n", cli.outputDirectory ? cli.outputDirectory : "(null)");
    printf(cli.runHydro ? "runHydro = True\n" : "runHydro = False\n");
    printf(cli.runTest ? "runTest = True\n" : "runTest = False\n");
    vah_config cfg;
    vah_config_load(cli.configDirectory ? cli.configDirectory : ".", &cfg);
    if (cli.runHydro) {
        int rc = vah_run(&cfg, cli.outputDirectory ? cli.outputDirectory : ".", cli.binary ? 1 : 0);
        if (rc != VAH_OK) {
            fprintf(stderr, "hydro failed (%d): %s\n", rc, vah_last_error());
            return 1;
        }
        printf("Done hydro.\n");
    }
    return 0;
}
