// sem_main.cpp -- the SEM proxy driver (src/sem/main/main.cpp:9-50) on
// libphc.so; with -receivers (and -fused) it also writes receivers.txt.
#include <cstdio>
#include <exception>

#include "SEMproxy.hpp"

int main(int argc, char *argv[])
{
    try {
        time_point<system_clock> startInitTime = system_clock::now();
        cout << "\n+================================= " << endl;
        cout << "| Initializing SEM Application ... " << endl;
        cout << "+================================= \n" << endl;

        SEMproxy semsim(argc, argv);
        semsim.initFiniteElem();

        cout << "\n+================================= " << endl;
        cout << "| Running SEM Application ...      " << endl;
        cout << "+================================= \n" << endl;

        time_point<system_clock> startRunTime = system_clock::now();
        semsim.run();

        if (!semsim.recNodes.empty() && !semsim.recTraces.empty()) {
            FILE *f = std::fopen("receivers.txt", "w");
            if (f) {
                const int n = semsim.info().myNumSamples;
                for (int s = 0; s < n; ++s) {
                    std::fprintf(f, "%d", s);
                    for (size_t r = 0; r < semsim.recNodes.size(); ++r)
                        std::fprintf(f, " %.7g", semsim.recTraces[r * n + s]);
                    std::fprintf(f, "\n");
                }
                std::fclose(f);
            }
        }
        cout << "\n+================================= " << endl;
        cout << "| SEM Application Finished.       " << endl;
        cout << "+================================= \n" << endl;
        cout << "Elapsed Initial Time : " << (startRunTime - startInitTime).count() / 1E9 << " seconds." << endl;
        cout << "Elapsed Compute Time : " << (system_clock::now() - startRunTime).count() / 1E9 << " seconds." << endl;
        cout << "Elapsed TotalExe Time : " << (system_clock::now() - startInitTime).count() / 1E9 << " seconds.\n"
             << endl;
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "sem_hip: %s\n", e.what());
        return 1;
    }
}
