// fd_main.cpp -- the FD proxy driver (src/fd/FDTDmain.cpp:9-108) on libphc.so.
// The per-step loop makes the reference's calls with the reference's
// arguments (FDTDmain.cpp:55-81); -fused replaces it with graph-captured runs
// of 50 steps.  Same prints and timers.
#include <algorithm>
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
        FDTDGRIDS myGrids;
        FDTDMODELS myModels;
        FDTDInit myInit;
        FDTDKernel myKernel;
        FDTDUtils myFDTDUtils;
        myInit.init_geometry(argc, argv, myGrids);
        myInit.init_coefficients(myGrids, myModels);
        myInit.init_source(myModels);
        myInit.init_models(myGrids, myModels);
        myInit.defineSpongeBoundary(myGrids, myModels.spongeArray);
        std::cout << "\n+================================= \n|  Running FDTD Application ...       \n"
                  << "+================================= \n\n";
        auto startRun = system_clock::now();
        duration<double> tRHS(0), tKernel(0), tOut(0);
        if (!myInit.fused) {
            for (int itSample = 0; itSample < myInit.nSamples; itSample++) {
                auto t0 = system_clock::now();
                myKernel.addRHS(myGrids, itSample, myModels, myInit.i2, myModels.pnGlobal);
                auto t1 = system_clock::now();
                if (myInit.usePML)
                    myKernel.computeOneStepPML(myGrids, myModels, myInit.i1, myInit.i2);
                else
                    myKernel.computeOneStepSB(myGrids, myModels, myInit.i1, myInit.i2);
                auto t2 = system_clock::now();
                std::swap(myInit.i1, myInit.i2);
                myFDTDUtils.output(myGrids, myModels.pnGlobal, itSample, myInit.i2, myInit.saveSnapShots);
                auto t3 = system_clock::now();
                tRHS += t1 - t0;
                tKernel += t2 - t1;
                tOut += t3 - t2;
            }
        } else {
            // graph-captured runs, each ending on the next output step
            phc_fd *h = myModels.device(myGrids);
            const int n = myInit.nSamples;
            int it = 0;
            while (it < n) {
                int last = ((it + 49) / 50) * 50;
                if (last > n - 1) last = n - 1;
                const int len = last - it + 1;
                auto t0 = system_clock::now();
                myModels.pnGlobal.push_host_changes();
                int cur = myInit.i2;
                phc_check(phc_fd_run(h, it, len, &cur), "phc_fd_run");
                phc_check(phc_fd_synchronize(h), "synchronize");
                myModels.pnGlobal.device_changed(0);
                myModels.pnGlobal.device_changed(1);
                auto t1 = system_clock::now();
                myInit.i2 = cur;
                myInit.i1 = 1 - cur;
                it += len;
                myFDTDUtils.output(myGrids, myModels.pnGlobal, last, myInit.i2, myInit.saveSnapShots);
                tKernel += t1 - t0;
                tOut += system_clock::now() - t1;
            }
        }
        if (myModels.dev) phc_check(phc_fd_synchronize(myModels.dev->h), "synchronize");
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
