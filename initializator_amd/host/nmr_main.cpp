// nmr_main.cpp -- the reference's driver (src/main.cpp:6-40) for one GPU:
//   nmr_exe gx gy gz [deviceShift] [--cpu-cadence] [--scheme upwind|maccormack]
//           [--print-config]
// argv[1..3] are the gradient-table indices, argv[4] shifts the device
// index (main.cpp:16-32).  --cpu-cadence runs solve() (results_cpu.csv)
// instead of solve_gpu(); --scheme overrides the update scheme of the loop
// (by default MacCormack for solve_gpu() and upwind for solve(), as in the
// reference); --print-config prints the parsed input as JSON and stops
// before any device work.  Reads data/ under the working directory.
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "NmrRelaxationSolve.h"
#include "nmr_hip.h"

int main(int argc, char** argv)
{
    int gradientNum[3] = {0, 0, 0};
    bool cpu = false, print = false;
    int pos = 0, deviceShift = 0, scheme = -1;
    for (int a = 1; a < argc; a++) {
        if (!std::strcmp(argv[a], "--cpu-cadence")) cpu = true;
        else if (!std::strcmp(argv[a], "--scheme") && a + 1 < argc) {
            const char* v = argv[++a];
            if (!std::strcmp(v, "upwind")) scheme = NMR_SCHEME_UPWIND;
            else if (!std::strcmp(v, "maccormack")) scheme = NMR_SCHEME_MACCORMACK;
            else {
                std::cerr << "unknown scheme " << v << std::endl;
                return 1;
            }
        }
        else if (!std::strcmp(argv[a], "--print-config")) print = true;
        else if (pos < 3) gradientNum[pos++] = std::atoi(argv[a]);
        else if (pos == 3) { deviceShift = std::atoi(argv[a]); pos++; }
    }
    if (pos < 3) std::cout << "Caught exception number:  0 gradient pulse failer!" << std::endl;
    std::cout << "Gradient pulses are:   X=" << gradientNum[0] << " Y=" << gradientNum[1] << " Z=" << gradientNum[2]
              << std::endl;
    NmrRelaxationSolve solver(1, 0, gradientNum);
    solver.set_scheme(scheme);
    if (print) {
        solver.read_config();
        std::cout << solver.config_json() << std::endl;
        return 0;
    }
    if (std::getenv("NMR_INIT_CONFIG")) { // full init, then the configuration (no device work)
        solver.init();
        std::cout << solver.config_json() << std::endl;
        return 0;
    }
    solver.init(); // files only: unsupported inputs stop here, before any device call
    int32_t ndev = 0;
    if (nmr_hip_device_count(&ndev) != NMR_OK || ndev < 1) {
        std::cerr << "no HIP device: " << nmr_hip_last_error() << std::endl;
        return 1;
    }
    std::cout << "HIP device " << deviceShift % ndev << std::endl;
    if (nmr_hip_set_device(deviceShift % ndev) != NMR_OK) {
        std::cerr << nmr_hip_last_error() << std::endl;
        return 1;
    }
    if (cpu) solver.solve();
    else solver.solve_gpu();
    return 0;
}
