// sem_main.cpp -- the SEM proxy driver (src/sem/main/main.cpp:9-50) on libphc.so.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <exception>
#include <iostream>

#include "SEM.hpp"

int main(int argc, char *argv[])
{
    try {
        auto t0 = std::chrono::system_clock::now();
        std::cout << "\n+================================= \n| Initializing SEM Application ... \n"
                  << "+================================= \n\n";
        SEMproxy semsim(argc, argv);
        semsim.initFiniteElem();
        std::cout << "\n+================================= \n| Running SEM Application ...      \n"
                  << "+================================= \n\n";
        auto t1 = std::chrono::system_clock::now();
        semsim.run();
        if (!semsim.recNodes.empty()) {
            FILE *f = std::fopen("receivers.txt", "w");
            if (f) {
                const int n = semsim.myInfo.myNumSamples;
                for (int s = 0; s < n; ++s) {
                    std::fprintf(f, "%d", s);
                    for (size_t r = 0; r < semsim.recNodes.size(); ++r)
                        std::fprintf(f, " %.7g", semsim.recTraces[r * n + s]);
                    std::fprintf(f, "\n");
                }
                std::fclose(f);
            }
        }
        std::cout << "\n+================================= \n| SEM Application Finished.       \n"
                  << "+================================= \n\n";
        auto t2 = std::chrono::system_clock::now();
        std::cout << "Elapsed Initial Time : " << std::chrono::duration<double>(t1 - t0).count() << " seconds." << std::endl;
        std::cout << "Elapsed Compute Time : " << std::chrono::duration<double>(t2 - t1).count() << " seconds." << std::endl;
        std::cout << "Elapsed TotalExe Time : " << std::chrono::duration<double>(t2 - t0).count() << " seconds.\n\n";
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "sem_hip: %s\n", e.what());
        return 1;
    }
}
