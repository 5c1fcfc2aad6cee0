// SEMproxy.cpp -- the SEM proxy driver (src/sem/main/SEMproxy.cpp:10-119) on
// the drop-in solver classes: the same calls with the same arguments, plus
// the run-time options of SEMproxy.hpp.
#include "SEMproxy.hpp"

#include <cmath>
#include <cstring>
#include <string>

SEMproxy::SEMproxy(int argc, char *argv[])
{
    int ex = (argc > 1) ? std::stoi(argv[1]) : 100;
    int ey = ex;
    int ez = ex;
    float lx = (argc > 2) ? std::stof(argv[2]) : 2000;
    float ly = lx;
    float lz = lx;
    bool gaussian = false, dim2 = false;
    unsigned long long seed = 0;
    parseOptions(argc, argv, gaussian, seed, dim2);
    SEMmesh simpleMesh{ex, ey, ez, lx, ly, lz, myInfo.myOrderNumber, dim2};
    myMesh = simpleMesh;
    if (gaussian)
        phc_check_sem(phc_sem_model_gaussian(myMesh.ex, dim2 ? 1 : myMesh.ey, myMesh.ez, 5.0, 2000.f, 4000.f, seed,
                                             myMesh.model.data()), "gaussian model");
}

void SEMproxy::parseOptions(int argc, char *argv[], bool &gaussian, unsigned long long &seed, bool &dim2)
{
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() {
            if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
            return std::string(argv[++i]);
        };
        if (a == "-order") myInfo.myOrderNumber = std::stoi(val());
        else if (a == "-dt") myInfo.myTimeStep = std::stof(val());
        else if (a == "-f0") myInfo.f0 = std::stof(val());
        else if (a == "-timemax") myInfo.myTimeMax = std::stof(val());
        else if (a == "-nsteps") nsteps = std::stoi(val());
        else if (a == "-element") mySolver.kernel = PHC_SEM_ELEMENT;
        else if (a == "-colors") mySolver.kernel = PHC_SEM_ELEMENT_COLOR;
        else if (a == "-gaussian") { gaussian = true; seed = std::stoull(val()); }
        else if (a == "-receivers") receivers = true;
        else if (a == "-fused") fused = true;
        else if (a == "-snapshots") mySolver.saveSnapShots = true;
        else if (a == "-2d") dim2 = true;
        else throw std::runtime_error("unknown argument " + a);
    }
    const int q = myInfo.myOrderNumber + 1;
    myInfo.nPointsPerElement = dim2 ? q * q : q * q * q;
    myInfo.myNumSamples = (int)(myInfo.myTimeMax / myInfo.myTimeStep);
    if (nsteps >= 0) myInfo.myNumSamples = nsteps;
}

// SEMproxy.cpp:25-38
void SEMproxy::initFiniteElem()
{
    getMeshInfo();
    init_arrays();
    init_source();
    mySolver.computeFEInit(myInfo, myMesh);
    if (receivers) {
        // 8 receivers on a line: x = lx/10 (r+1), y = ly/2, z = 100 m (nearest nodes)
        recNodes.clear();
        for (int r = 0; r < 8; ++r) {
            const double x = myMesh.lx / 10.0 * (r + 1), y = myMesh.ly / 2.0, z = 100.0;
            const int ix = (int)std::lround(x / (myMesh.lx / (double)(myMesh.nx - 1)));
            const int iy = (myMesh.ny == 1) ? 0 : (int)std::lround(y / (myMesh.ly / (double)(myMesh.ny - 1)));
            const int iz = (int)std::lround(z / (myMesh.lz / (double)(myMesh.nz - 1)));
            recNodes.push_back(myMesh.node(ix, iy, iz));
        }
        mySolver.setReceivers((int)recNodes.size(), recNodes.data());
    }
}

// SEMproxy.cpp:41-64
void SEMproxy::run()
{
    if (fused) {
        runFused();
        return;
    }
    duration<double> totalComputeTime(0), totalOutputTime(0);
    for (int indexTimeSample = 0; indexTimeSample < myInfo.myNumSamples; indexTimeSample++) {
        auto startComputeTime = system_clock::now();
        mySolver.computeOneStep(indexTimeSample, myInfo.myOrderNumber, myInfo.nPointsPerElement, i1, i2, myInfo,
                                myRHSTerm, pnGlobal, rhsElement);
        totalComputeTime += system_clock::now() - startComputeTime;

        auto startOutputTime = system_clock::now();
        mySolver.outputPnValues(myMesh, indexTimeSample, i1, myInfo.myElementSource, pnGlobal);
        swap(i1, i2);
        totalOutputTime += system_clock::now() - startOutputTime;
    }
    mySolver.synchronize();
    cout << "------------------------------------------------ " << endl;
    cout << "\n---- Elapsed Kernel Time : " << totalComputeTime.count() << " seconds." << endl;
    cout << "---- Elapsed Output Time : " << totalOutputTime.count() << " seconds." << endl;
    cout << "------------------------------------------------ " << endl;
}

// Graph-captured runs of up to 100 steps, each ending on an output step;
// receivers and the source trace are recorded on the device.
void SEMproxy::runFused()
{
    const int n = myInfo.myNumSamples;
    std::vector<float> samples((size_t)n);
    for (int j = 0; j < n; ++j) samples[(size_t)j] = myRHSTerm(0, j);
    mySolver.setSource(rhsElement[0], samples.data(), n);
    recTraces.assign(recNodes.size() * (size_t)n, 0.f);
    std::vector<float> part;
    auto t0 = system_clock::now();
    int it = 0;
    while (it < n) {
        int last = ((it + 99) / 100) * 100;
        if (last > n - 1) last = n - 1;
        const int len = last - it + 1;
        int a = i1, b = i2;
        mySolver.run(pnGlobal, it, len, &a, &b);
        if (!recNodes.empty()) {
            part.assign(recNodes.size() * (size_t)len, 0.f);
            mySolver.traces(nullptr, part.data());
            for (size_t r = 0; r < recNodes.size(); ++r)
                std::memcpy(&recTraces[r * n + it], &part[r * len], sizeof(float) * len);
        }
        // after computeOneStep of step `last` and before its swap, the
        // reference prints level i1 (SEMproxy.cpp:48-53)
        std::swap(a, b);
        i1 = a;
        i2 = b;
        mySolver.outputPnValues(myMesh, last, i1, myInfo.myElementSource, pnGlobal);
        std::swap(i1, i2);
        it += len;
    }
    mySolver.synchronize();
    const double s = duration<double>(system_clock::now() - t0).count();
    printf("------------------------------------------------ \n\n---- Elapsed Kernel Time : %g seconds.\n", s);
}

// SEMproxy.cpp:67-81
void SEMproxy::getMeshInfo()
{
    myInfo.numberOfNodes = myMesh.getNumberOfNodes();
    myInfo.numberOfElements = myMesh.getNumberOfElements();
    myInfo.numberOfPointsPerElement = myMesh.getNumberOfPointsPerElement();
    myInfo.numberOfInteriorNodes = myMesh.getNumberOfInteriorNodes();
    myInfo.numberOfBoundaryNodes = myMesh.getNumberOfBoundaryNodes();
    myInfo.numberOfBoundaryFaces = myMesh.getNumberOfBoundaryFaces();
}

// SEMproxy.cpp:84-90
void SEMproxy::init_arrays()
{
    cout << "Allocate host memory for source and pressure values ..." << endl;
    myRHSTerm = allocateArray2D<arrayReal>(myInfo.myNumberOfRHS, myInfo.myNumSamples, "RHSTerm");
    rhsElement = allocateVector<vectorInt>(myInfo.myNumberOfRHS, "rhsElement");
    pnGlobal = allocateArray2D<arrayReal>(myInfo.numberOfNodes, 2, "pnGlobal");
}

// SEMproxy.cpp:93-119: source at (1001, 1001, 1001)
void SEMproxy::init_source()
{
    arrayReal myRHSLocation = allocateArray2D<arrayReal>(myInfo.myNumberOfRHS, 3, "RHSLocation");
    myRHSLocation(0, 0) = 1001;
    myRHSLocation(0, 1) = 1001;
    myRHSLocation(0, 2) = 1001;
    cout << "\nSource location: " << myRHSLocation(0, 0) << ", " << myRHSLocation(0, 1) << ", "
         << myRHSLocation(0, 2) << endl;
    for (int i = 0; i < myInfo.myNumberOfRHS; i++)
        rhsElement[i] = myMesh.getElementNumberFromPoints(myRHSLocation(i, 0), myRHSLocation(i, 1),
                                                          myRHSLocation(i, 2));
    vector<float> sourceTerm = myUtils.computeSourceTerm(myInfo.myNumSamples, myInfo.myTimeStep, myInfo.f0,
                                                         myInfo.sourceOrder);
    for (int j = 0; j < myInfo.myNumSamples; j++) {
        myRHSTerm(0, j) = sourceTerm[j];
        if (j % 100 == 0) cout << "Sample " << j << "\t: sourceTerm = " << sourceTerm[j] << endl;
    }
    myInfo.myElementSource = rhsElement[0];
    cout << "Element number for the source location: " << myInfo.myElementSource << endl << endl;
}
