// SEMproxy.hpp -- the SEM proxy driver class (src/sem/main/SEMproxy.hpp:13-66)
// over the drop-in solver classes of SEM.hpp.  Same members and methods as
// the reference, so its SEMproxy.cpp and main.cpp compile against this header
// unchanged; this library's own SEMproxy.cpp adds run-time options.  Like the
// reference's headers (SEMmesh.hpp:6, utils.hpp:6) it puts std and
// std::chrono in scope for the code that includes it.
#pragma once

#include <chrono>
#include <iostream>
#include <utility>
#include <vector>

#include "SEM.hpp"

using namespace std;
using namespace std::chrono;

class SEMproxy {
  public:
    // SEMproxy.cpp:10-22: argv[1] = ex (= ey = ez), argv[2] = lx (= ly = lz)
    SEMproxy(int argc, char *argv[]);
    ~SEMproxy(){};

    void initFiniteElem();
    void run();
    void init_source();
    void init_arrays();
    void getMeshInfo();

    // ---- additions (no reference counterpart): options after argv[2]
    //   -order P, -dt DT, -f0 F, -timemax T, -nsteps N, -element (element
    //   kernel), -colors (element kernel, 8-colour scatter), -gaussian SEED
    //   (smooth random model 2000-4000 m/s), -receivers (8 on the line
    //   y = ly/2, z = 100 m, fused runs), -fused (graph-captured 100-step
    //   runs), -snapshots (ASCII snapshot every 100 steps), -2d (SEM2D)
    int nsteps = -1;
    bool fused = false, receivers = false;
    std::vector<int> recNodes;
    std::vector<float> recTraces;  // receiver-major, myNumSamples per receiver
    const SEMinfo &info() const { return myInfo; }

  private:
    int i1 = 0;
    int i2 = 1;

    SEMinfo myInfo;
    SEMmesh myMesh;

    SEMsolver mySolver;
    SolverUtils myUtils;

    arrayReal myRHSTerm;
    arrayReal pnGlobal;
    vectorInt rhsElement;

    void parseOptions(int argc, char *argv[], bool &gaussian, unsigned long long &seed, bool &dim2);
    void runFused();
};
