// gfm2d: settings-file driver (main.cpp:8-43 of the reference): reads the
// settings file named by argv[1] and runs the onefluid or twofluid
// simulation on the GPU.  `gfm2d --parse FILE` prints the parsed settings.
#include <cstdio>
#include <cstring>
#include <memory>

#include "gfm2d.hpp"

int main(int argc, char* argv[]) {
    using namespace gfm2d;
    if (argc < 2) {
        std::fprintf(stderr, "usage: gfm2d SETTINGSFILE | gfm2d --parse SETTINGSFILE\n");
        return 2;
    }
    try {
        if (std::strcmp(argv[1], "--parse") == 0 && argc > 2) {
            GFM_settingsfile SF;
            SF.read_settings_file(argv[2]);
            std::printf("Nx %d\nNy %d\nCFL %.17g\nsim_type %s\nFS %s\ntest_case %s\nITM %s\nGFM %s\noutputpath %s\nbasename %s\n",
                        SF.Nx, SF.Ny, SF.CFL, SF.sim_type.c_str(), SF.FS.c_str(), SF.test_case.c_str(), SF.ITM.c_str(),
                        SF.GFM.c_str(), SF.outputpath.c_str(), SF.basename.c_str());
            return 0;
        }
        std::printf("Beginning 2D GFM code..\n");
        GFM_settingsfile SF;
        SF.read_settings_file(argv[1]);
        std::printf("Settings file loaded.\nSimulation name is %s\n", SF.basename.c_str());
        std::shared_ptr<sim_base> sim;
        if (SF.sim_type == "onefluid") sim = std::make_shared<sim_onefluid>();
        else if (SF.sim_type == "twofluid") sim = std::make_shared<sim_twofluid>();
        else {
            std::fprintf(stderr, "Invalid sim type %s\n", SF.sim_type.c_str());
            return 1;
        }
        sim->run_sim(SF);
        std::printf("Simulation complete!\n");
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "gfm2d: %s\n", ex.what());
        return 1;
    }
    return 0;
}
