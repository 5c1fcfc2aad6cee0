// SEM.hpp -- C++ drop-in for the reference SEM proxy's classes, backed by
// libphc.so (the wavefield lives on the GPU; pnGlobal is a host mirror that is
// refreshed only for output).
//
//   SEMinfo                   src/sem/main/SEMdata.hpp:3-30
//   SEMmesh                   src/sem/mesh/SEMmesh.hpp:9-118
//   SEMsolver                 src/sem/solver/SEMsolver.hpp:18-91
//   SEMproxy                  src/sem/main/SEMproxy.hpp, SEMproxy.cpp:10-119
//
// Same class and method names, argument meaning and argv (argv[1] = ex,
// argv[2] = lx).  The order, time step, source frequency and duration are
// compile-time constants in the reference (SEMdata.hpp:13-22); here they are
// runtime options with the reference's values as defaults:
//   -order P, -dt DT, -f0 F, -timemax T, -nsteps N, -element (element kernel),
//   -gaussian SEED (smooth random model 2000-4000 m/s), -receivers (8 on the
//   line y = ly/2, z = 100 m), -fused (graph-captured runs), -snapshots (the
//   reference's SEM_SAVE_SNAPSHOTS build option: ASCII snapshot every 100 steps),
//   -2d (the reference's SEM2D build option: ey = 0, quadrilaterals on the x-z
//   plane, paraxial absorbing edges; SEMmacros.hpp:7-11, SEMsolver.cpp:91-123).
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "phc.h"
#include "phc_setup.h"

inline void phc_check_sem(phc_status s, const char *what)
{
    if (s != PHC_OK) throw std::runtime_error(std::string(what) + ": " + phc_last_error());
}

// Host arrays with the reference's container interface (dataType.hpp:14-80).
template <class T> struct Vector {
    std::vector<T> data;
    Vector(int n = 0) : data(n) {}
    T &operator[](int i) { return data[i]; }
    const T &operator[](int i) const { return data[i]; }
    int size() const { return (int)data.size(); }
};
template <class T> struct Array2D {
    std::vector<T> data;
    int rows = 0, cols = 0;
    Array2D(int r = 0, int c = 0) : data((size_t)r * c, T(0)), rows(r), cols(c) {}
    T &operator()(int r, int c) { return data[(size_t)r * cols + c]; }
    const T &operator()(int r, int c) const { return data[(size_t)r * cols + c]; }
};
using vectorReal = Vector<float>;
using vectorInt = Vector<int>;
using arrayReal = Array2D<float>;

struct SEMinfo {
    int numberOfNodes = 0, numberOfElements = 0, numberOfPointsPerElement = 0;
    int numberOfInteriorNodes = 0;
    int myNumberOfRHS = 1;
    int myOrderNumber = 2;
    float myTimeStep = 0.001f;
    float f0 = 10.f;
    float myTimeMax = 0.01f;
    int sourceOrder = 1;
    int myNumSamples = 0;
    int myElementSource = 0;
};

class SEMmesh {
  public:
    int ex = 0, ey = 0, ez = 0, nx = 0, ny = 0, nz = 0, order = 0;
    float lx = 0, ly = 0, lz = 0, hx = 0, hy = 0, hz = 0;
    std::vector<float> model;
    SEMmesh() = default;
    // SEMmesh.cpp:3-26; dim2 is the reference's SEM2D build, which zeroes ey
    // and ly (ZEROED2D, SEMmacros.hpp:10) and keeps one node layer (ny = 1)
    SEMmesh(const int &ex_in, const int &ey_in, const int &ez_in, const float &lx_in,
            const float &ly_in, const float &lz_in, const int &order_in, bool dim2 = false)
        : ex(ex_in), ey(dim2 ? 0 : ey_in), ez(ez_in), order(order_in), lx(lx_in), ly(dim2 ? 0.f : ly_in),
          lz(lz_in)
    {
        nx = ex * order + 1;
        ny = (ey == 0) ? 1 : ey * order + 1;
        nz = ez * order + 1;
        hx = (float)(lx / (1. * ex));
        hy = (ey == 0) ? 1.f : (float)(ly / (1. * ey));
        hz = (float)(lz / (1. * ez));
        model.assign((size_t)ex * (ey == 0 ? 1 : ey) * ez, 1500.f);  // SEMmesh::getModel, SEMmesh.cpp:288-312
        printf("SEMmesh initiliazed\nex, ey, ez=%d, %d, %d\nnx, ny, nz=%d, %d, %d\n", ex, ey, ez, nx, ny, nz);
    }
    int getNumberOfNodes() const { return nx * ny * nz; }
    int getNumberOfElements() const { return (ey == 0) ? ex * ez : ex * ey * ez; }
    int getNumberOfPointsPerElement() const
    {
        return (ny == 1) ? (order + 1) * (order + 1) : (order + 1) * (order + 1) * (order + 1);
    }
    int getNumberOfInteriorNodes() const { return (ny == 1) ? (nx - 2) * (nz - 2) : (nx - 2) * (ny - 2) * (nz - 2); }
    int getNx() const { return nx; }
    int getNy() const { return ny; }
    int getNz() const { return nz; }
    int getElementNumberFromPoints(const float &x, const float &y, const float &z) const
    {
        return phc_sem_element_from_point(ex, ey, ez, lx, ly, lz, x, y, z);
    }
    // node id of the grid point (ix, iy, iz) (SEMmesh.cpp:214-223 numbering)
    int node(int ix, int iy, int iz) const { return ix + iz * nx + iy * nx * nz; }
    // SEMmesh.cpp:69-79: the element sizes, returned as int
    int getDx() const { return (int)hx; }
    int getDy() const { return (int)hy; }
    int getDz() const { return (int)hz; }
    // SEMmesh.cpp:624-654: the node plane iy = ny/2 - 1 of level i1 as ASCII
    // "x z value" lines (x = ix * element size, z = iz * element size, as the
    // reference writes them), file "snapshot<step>"
    template <class Field> void saveSnapShot(const int indexTimeStep, const int i1, const Field &u) const
    {
        const float dx = getDx(), dz = getDz();
        const int numberOfNodes = nx * nz;
        const int offset = (ny == 1) ? 0 : nx * nz * (ny / 2 - 1);
        std::ofstream snapFile("snapshot" + std::to_string(indexTimeStep), std::ios::out | std::ios::trunc);
        if (!snapFile) throw std::runtime_error("cannot open snapshot file");
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < nz; j++) {
                (void)numberOfNodes;
                snapFile << i * dx << " " << j * dz << " " << u(offset + i + j * nx, i1) << std::endl;
            }
    }
};

class SEMsolver {
  public:
    phc_sem *dev = nullptr;      // 3D mesh
    phc_sem2d *dev2 = nullptr;   // SEM2D mesh (ey = 0)
    int kernel = PHC_SEM_SEPARABLE;
    std::vector<float> rhs;
    ~SEMsolver()
    {
        if (dev) phc_sem_destroy(dev);
        if (dev2) phc_sem2d_destroy(dev2);
    }
    phc_status getField(int level, float *host) const
    {
        return dev2 ? phc_sem2d_get_field(dev2, level, host) : phc_sem_get_field(dev, level, host);
    }
    phc_status run(int it0, int n, int *a, int *b)
    {
        return dev2 ? phc_sem2d_run(dev2, it0, n, a, b) : phc_sem_run(dev, it0, n, a, b);
    }
    phc_status traces(float *src, float *rec)
    {
        return dev2 ? phc_sem2d_get_traces(dev2, src, rec) : phc_sem_get_traces(dev, src, rec);
    }
    phc_status setReceivers(int n, const int *nodes)
    {
        return dev2 ? phc_sem2d_set_receivers(dev2, n, nodes) : phc_sem_set_receivers(dev, n, nodes);
    }
    phc_status synchronize() { return dev2 ? phc_sem2d_synchronize(dev2) : phc_sem_synchronize(dev); }
    // SEMsolver.cpp:12-17: mesh, basis, geometry, mass -- built once in FP64
    void computeFEInit(SEMinfo &info, const SEMmesh &mesh, const arrayReal &RHS)
    {
        rhs.assign(RHS.data.begin(), RHS.data.begin() + info.myNumSamples);
        phc_sem_params p;
        std::memset(&p, 0, sizeof(p));
        p.ex = mesh.ex; p.ey = mesh.ey; p.ez = mesh.ez;
        p.lx = mesh.lx; p.ly = mesh.ly; p.lz = mesh.lz;
        p.order = info.myOrderNumber;
        p.model = mesh.model.data();
        p.dt = info.myTimeStep;
        p.rhs = rhs.data();
        p.nsamples = info.myNumSamples;
        p.src_element = info.myElementSource;
        p.kernel = kernel;
        p.weights_ref = 1;
        if (dev) phc_sem_destroy(dev);
        if (dev2) phc_sem2d_destroy(dev2);
        dev = nullptr;
        dev2 = nullptr;
        if (mesh.ey == 0) phc_check_sem(phc_sem2d_create(&p, &dev2), "phc_sem2d_create");
        else phc_check_sem(phc_sem_create(&p, &dev), "phc_sem_create");
    }
    // SEMsolver.cpp:21-125 (same signature; the source samples and the source
    // element were handed over at computeFEInit, pnGlobal stays on the GPU)
    void computeOneStep(const int &timeSample, const int &, const int &, const int &i1, const int &i2,
                        SEMinfo &, const arrayReal &, arrayReal const &, const vectorInt &)
    {
        phc_check_sem(dev2 ? phc_sem2d_compute_one_step(dev2, timeSample, i1, i2)
                           : phc_sem_compute_one_step(dev, timeSample, i1, i2), "computeOneStep");
    }
    // SEMsolver.cpp:127-142
    void outputPnValues(const SEMmesh &mesh, const int &indexTimeStep, int &i1, int &myElementSource,
                        arrayReal &pnGlobal)
    {
        if (indexTimeStep % 100 == 0) {
            std::vector<float> f((size_t)mesh.getNumberOfNodes());
            phc_check_sem(getField(i1, f.data()), "get_field");
            for (size_t n = 0; n < f.size(); ++n) pnGlobal((int)n, i1) = f[n];
            const int e = myElementSource;
            const int node = (e % mesh.ex) * mesh.order + ((e / mesh.ex) % mesh.ez) * mesh.order * mesh.nx +
                             (mesh.ey == 0 ? 0 : (e / (mesh.ex * mesh.ez)) * mesh.order * mesh.nx * mesh.nz);
            printf("TimeStep=%d;  pnGlobal @ elementSource location %d after computeOneStep = %g\n",
                   indexTimeStep, myElementSource, pnGlobal(node, i1));
            if (saveSnapShots) mesh.saveSnapShot(indexTimeStep, i1, pnGlobal);
        }
    }
    bool saveSnapShots = false;  // the reference's SEM_SAVE_SNAPSHOTS build option
};

class SEMproxy {
  public:
    SEMinfo myInfo;
    SEMmesh myMesh;
    SEMsolver mySolver;
    arrayReal myRHSTerm, pnGlobal;
    vectorInt rhsElement;
    int i1 = 0, i2 = 1;
    int nsteps = -1;
    bool fused = false, receivers = false;
    std::vector<int> recNodes;
    std::vector<float> recTraces;

    // SEMproxy.cpp:10-22 (argv[1] = ex, argv[2] = lx), then the options
    SEMproxy(int argc, char *argv[])
    {
        const int ex = (argc > 1) ? std::stoi(argv[1]) : 100;
        const float lx = (argc > 2) ? std::stof(argv[2]) : 2000.f;
        unsigned long long seed = 0;
        bool gaussian = false, dim2 = false;
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
            else if (a == "-gaussian") { gaussian = true; seed = std::stoull(val()); }
            else if (a == "-receivers") receivers = true;
            else if (a == "-fused") fused = true;
            else if (a == "-snapshots") mySolver.saveSnapShots = true;
            else if (a == "-2d") dim2 = true;
            else throw std::runtime_error("unknown argument " + a);
        }
        myMesh = SEMmesh(ex, ex, ex, lx, lx, lx, myInfo.myOrderNumber, dim2);
        if (gaussian)
            phc_check_sem(phc_sem_model_gaussian(ex, dim2 ? 1 : ex, ex, 5.0, 2000.f, 4000.f, seed,
                                                 myMesh.model.data()), "gaussian model");
    }

    // SEMproxy.cpp:25-38
    void initFiniteElem()
    {
        myInfo.numberOfNodes = myMesh.getNumberOfNodes();
        myInfo.numberOfElements = myMesh.getNumberOfElements();
        myInfo.numberOfPointsPerElement = myMesh.getNumberOfPointsPerElement();
        myInfo.numberOfInteriorNodes = myMesh.getNumberOfInteriorNodes();
        myInfo.myNumSamples = (int)(myInfo.myTimeMax / myInfo.myTimeStep);
        if (nsteps >= 0) myInfo.myNumSamples = nsteps;
        printf("Number of nodes: %d\nNumber of elements: %d\n", myInfo.numberOfNodes, myInfo.numberOfElements);
        // init_arrays (SEMproxy.cpp:84-90)
        myRHSTerm = arrayReal(myInfo.myNumberOfRHS, myInfo.myNumSamples);
        rhsElement = vectorInt(myInfo.myNumberOfRHS);
        pnGlobal = arrayReal(myInfo.numberOfNodes, 2);
        // init_source (SEMproxy.cpp:93-119): source at (1001, 1001, 1001)
        printf("\nSource location: 1001, 1001, 1001\n");
        rhsElement[0] = myMesh.getElementNumberFromPoints(1001.f, 1001.f, 1001.f);
        std::vector<float> st(myInfo.myNumSamples);
        phc_check_sem(phc_source_term(myInfo.myNumSamples, myInfo.myTimeStep, myInfo.f0, myInfo.sourceOrder,
                                      PHC_EXP_FLOAT, st.data()), "source term");
        for (int j = 0; j < myInfo.myNumSamples; ++j) {
            myRHSTerm(0, j) = st[j];
            if (j % 100 == 0) printf("Sample %d\t: sourceTerm = %g\n", j, st[j]);
        }
        myInfo.myElementSource = rhsElement[0];
        printf("Element number for the source location: %d\n\n", myInfo.myElementSource);
        mySolver.computeFEInit(myInfo, myMesh, myRHSTerm);
        if (receivers) {
            // 8 receivers on a line: x = ly/10 (r+1), y = ly/2, z = 100 m (nearest nodes)
            recNodes.clear();
            const int p = myMesh.order;
            for (int r = 0; r < 8; ++r) {
                const double x = myMesh.lx / 10.0 * (r + 1), y = myMesh.ly / 2.0, z = 100.0;
                const int ix = (int)std::lround(x / (myMesh.lx / (double)(myMesh.nx - 1)));
                const int iy = (myMesh.ny == 1) ? 0 : (int)std::lround(y / (myMesh.ly / (double)(myMesh.ny - 1)));
                const int iz = (int)std::lround(z / (myMesh.lz / (double)(myMesh.nz - 1)));
                (void)p;
                recNodes.push_back(myMesh.node(ix, iy, iz));
            }
            phc_check_sem(mySolver.setReceivers((int)recNodes.size(), recNodes.data()), "receivers");
        }
    }

    // SEMproxy.cpp:41-64
    double run()
    {
        auto t0 = std::chrono::system_clock::now();
        if (!fused) {
            for (int it = 0; it < myInfo.myNumSamples; ++it) {
                mySolver.computeOneStep(it, myInfo.myOrderNumber, myInfo.numberOfPointsPerElement, i1, i2, myInfo,
                                        myRHSTerm, pnGlobal, rhsElement);
                mySolver.outputPnValues(myMesh, it, i1, myInfo.myElementSource, pnGlobal);
                std::swap(i1, i2);
            }
        } else {
            const int n = myInfo.myNumSamples;
            recTraces.assign(recNodes.size() * (size_t)n, 0.f);
            std::vector<float> part;
            int it = 0;
            while (it < n) {
                int last = ((it + 99) / 100) * 100;
                if (last > n - 1) last = n - 1;
                const int len = last - it + 1;
                int a = i1, b = i2;
                phc_check_sem(mySolver.run(it, len, &a, &b), "run");
                if (!recNodes.empty()) {
                    part.assign(recNodes.size() * (size_t)len, 0.f);
                    phc_check_sem(mySolver.traces(nullptr, part.data()), "traces");
                    for (size_t r = 0; r < recNodes.size(); ++r)
                        std::memcpy(&recTraces[r * n + it], &part[r * len], sizeof(float) * len);
                }
                // after computeOneStep of step `last` and before its swap, the
                // reference prints level i1 (SEMproxy.cpp:48-53)
                std::swap(a, b);
                i1 = a; i2 = b;
                mySolver.outputPnValues(myMesh, last, i1, myInfo.myElementSource, pnGlobal);
                std::swap(i1, i2);
                it += len;
            }
        }
        phc_check_sem(mySolver.synchronize(), "synchronize");
        const double s = std::chrono::duration<double>(std::chrono::system_clock::now() - t0).count();
        printf("------------------------------------------------ \n\n---- Elapsed Kernel Time : %g seconds.\n", s);
        return s;
    }
};
