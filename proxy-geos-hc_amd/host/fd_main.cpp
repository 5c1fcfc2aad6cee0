// fd_main.cpp -- the FD proxy driver (src/fd/FDTDmain.cpp:9-108) on libphc.so.
// Same argv, same prints, same timers; the step runs on the GPU.
#include <chrono>
#include <cstdio>
#include <exception>
#include <iostream>

#include "FDTD.hpp"

using namespace std::chrono;

int main(int argc, char *argv[])
{
    try {
        auto startInit = system_clock::now();
        std::cout << "\n+==================================== \n| Initializing FDTD Application ...       \n"
                  << "+==================================== \n\n";
        FDTDGRIDS g;
        FDTDMODELS m;
        FDTDInit init;
        FDTDKernel kernel;
        FDTDUtils utils;
        init.init_geometry(argc, argv, g);
        init.init_coefficients(g, m);
        init.init_source(m);
        init.init_models(g, m);
        init.defineSpongeBoundary(g, m);
        std::cout << "\n+================================= \n|  Running FDTD Application ...       \n"
                  << "+================================= \n\n";
        const int n = init.nsteps >= 0 ? std::min(init.nsteps, init.nSamples) : init.nSamples;
        auto startRun = system_clock::now();
        duration<double> tRHS(0), tKernel(0), tOut(0);
        if (!init.fused) {
            for (int it = 0; it < n; ++it) {
                auto t0 = system_clock::now();
                kernel.addRHS(g, it, m, init.i2);
                auto t1 = system_clock::now();
                if (init.usePML)
                    kernel.computeOneStepPML(g, m, init.i1, init.i2);
                else
                    kernel.computeOneStepSB(g, m, init.i1, init.i2);
                auto t2 = system_clock::now();
                std::swap(init.i1, init.i2);
                utils.output(g, m, it, init.i2, init.saveSnapShots);
                auto t3 = system_clock::now();
                tRHS += t1 - t0;
                tKernel += t2 - t1;
                tOut += t3 - t2;
            }
        } else {
            // graph-captured runs, each ending on the next output step
            int it = 0;
            while (it < n) {
                int last = ((it + 49) / 50) * 50;
                if (last > n - 1) last = n - 1;
                const int len = last - it + 1;
                auto t0 = system_clock::now();
                int cur = init.i2;
                phc_check(phc_fd_run(m.dev, it, len, &cur), "phc_fd_run");
                phc_check(phc_fd_synchronize(m.dev), "synchronize");
                auto t1 = system_clock::now();
                init.i2 = cur;
                init.i1 = 1 - cur;
                it += len;
                utils.output(g, m, last, init.i2, init.saveSnapShots);
                tKernel += t1 - t0;
                tOut += system_clock::now() - t1;
            }
        }
        phc_check(phc_fd_synchronize(m.dev), "synchronize");
        std::cout << "\n+================================= \n|  FDTD Application Finished.       \n"
                  << "+================================= \n\n";
        std::cout << "Elapsed Initialization Time : " << duration<double>(startRun - startInit).count()
                  << " seconds.\n\n";
        std::cout << "Elapsed ComputeLoop RunTime : " << duration<double>(system_clock::now() - startRun).count()
                  << " seconds." << std::endl;
        std::cout << "------------------------------------------------ " << std::endl;
        std::cout << "---- Elapsed Kernel Time : " << tKernel.count() << " seconds." << std::endl;
        std::cout << "---- Elapsed AddRHS Time : " << tRHS.count() << " seconds." << std::endl;
        std::cout << "---- Elapsed Output Time : " << tOut.count() << " seconds." << std::endl;
        std::cout << "------------------------------------------------ " << std::endl;
        std::cout << "Elapsed TotalExecution Time : " << duration<double>(system_clock::now() - startInit).count()
                  << " seconds.\n\n";
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "fd_hip: %s\n", e.what());
        return 1;
    }
}
