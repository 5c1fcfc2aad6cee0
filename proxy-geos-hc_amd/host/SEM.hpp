// SEM.hpp -- C++ drop-in for the reference SEM proxy's solver classes, backed
// by libphc.so (the wavefield levels live on the GPU).
//
//   SEMinfo                   src/sem/main/SEMdata.hpp:3-30
//   SEMmesh                   src/sem/mesh/SEMmesh.hpp:9-118
//   SEMsolver                 src/sem/solver/SEMsolver.hpp:18-91
//
// Same class names, members and method signatures, so the reference's
// SEMproxy (SEMproxy.cpp:25-64) calls them unchanged:
//   computeFEInit(myInfo, myMesh);
//   computeOneStep(it, order, nPointsPerElement, i1, i2, myInfo, myRHSTerm,
//                  pnGlobal, rhsElement);
//   outputPnValues(myMesh, it, i1, myInfo.myElementSource, pnGlobal).
// computeOneStep injects myRHSTerm(r, it) at rhsElement[r] as the caller holds
// them at that call (SEMsolver.cpp:37-42), then steps.  The pnGlobal it is
// given is bound to the device's levels at the first call: reads after a
// step refresh the host copy, host writes are sent back before the next
// step (phc_arrays.hpp).
//
// The order, time step, frequency and duration are compile-time constants in
// the reference (SEMdata.hpp:13-22); here they are SEMinfo members with the
// reference's values as defaults, so a driver can change them at run time.
// SEM2D (ey = 0, quadrilaterals on the x-z plane, paraxial edges) is the
// reference's SEM2D build option (SEMmacros.hpp:7-11, SEMsolver.cpp:91-123).
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "phc.h"
#include "phc_arrays.hpp"
#include "phc_setup.h"
#include "phc_solver_utils.hpp"

inline void phc_check_sem(phc_status s, const char *what)
{
    if (s != PHC_OK) throw std::runtime_error(std::string(what) + ": " + phc_last_error());
}

struct SEMinfo {
    // from the mesh (SEMproxy::getMeshInfo)
    int numberOfNodes = 0;
    int numberOfElements = 0;
    int numberOfPointsPerElement = 0;
    int numberOfInteriorNodes = 0;
    int numberOfBoundaryNodes = 0;
    int numberOfBoundaryFaces = 0;
    // SEMdata.hpp:13-22 (constants there)
    int myNumberOfRHS = 1;
    int myOrderNumber = 2;
    float myTimeStep = 0.001f;
    int nPointsPerElement = (myOrderNumber + 1) * (myOrderNumber + 1) * (myOrderNumber + 1);
    float f0 = 10.f;
    float myTimeMax = 0.01f;
    int sourceOrder = 1;
    int myNumSamples = (int)(myTimeMax / myTimeStep);
    int myElementSource = 0;
};

class SEMmesh {
  public:
    int ex = 0, ey = 0, ez = 0, nx = 0, ny = 0, nz = 0, order = 0;
    float lx = 0, ly = 0, lz = 0, hx = 0, hy = 0, hz = 0;
    std::vector<float> model;  // wave speed per element (SEMmesh::getModel)
    SEMmesh() = default;
    // SEMmesh.cpp:3-26; dim2 is the reference's SEM2D build, which zeroes ey
    // and ly (ZEROED2D, SEMmacros.hpp:10) and keeps one node layer (ny = 1)
    SEMmesh(const int &ex_in, const int &ey_in, const int &ez_in, const float &lx_in, const float &ly_in,
            const float &lz_in, const int &order_in, bool dim2 = false)
        : ex(ex_in), ey(dim2 ? 0 : ey_in), ez(ez_in), order(order_in), lx(lx_in), ly(dim2 ? 0.f : ly_in), lz(lz_in)
    {
        nx = ex * order + 1;
        ny = (ey == 0) ? 1 : ey * order + 1;
        nz = ez * order + 1;
        hx = (float)(lx / (1. * ex));
        hy = (ey == 0) ? 1.f : (float)(ly / (1. * ey));
        hz = (float)(lz / (1. * ez));
        model.assign((size_t)ex * (ey == 0 ? 1 : ey) * ez, 1500.f);  // SEMmesh.cpp:288-312
        printf("SEMmesh initiliazed\nex, ey, ez=%d, %d, %d\nnx, ny, nz=%d, %d, %d\n", ex, ey, ez, nx, ny, nz);
    }
    int getNumberOfNodes() const { return nx * ny * nz; }
    int getNumberOfElements() const { return (ey == 0) ? ex * ez : ex * ey * ez; }
    int getNumberOfPointsPerElement() const
    {
        return (ny == 1) ? (order + 1) * (order + 1) : (order + 1) * (order + 1) * (order + 1);
    }
    int getNumberOfInteriorElements() const
    {
        return (ey == 0) ? (ex - 2) * (ez - 2) : (ex - 2) * (ey - 2) * (ez - 2);
    }
    int getNumberOfInteriorNodes() const { return (ny == 1) ? (nx - 2) * (nz - 2) : (nx - 2) * (ny - 2) * (nz - 2); }
    // SEMmesh.cpp:433-438 (the 3D count is the reference's formula as written)
    int getNumberOfBoundaryFaces() const { return (ey == 0) ? 2 * (ex + ez) : 2 * (ex * ey + ey * ez + ex * ez); }
    int getNumberOfBoundaryNodes() const
    {
        return (ey == 0) ? 2 * (nx + nz) - 4 : 2 * ((nx - 1) * (nz - 1) + (nx - 1) * (ny - 1) + (ny - 1) * (ny - 1));
    }
    int getNx() const { return nx; }
    int getNy() const { return ny; }
    int getNz() const { return nz; }
    // SEMmesh.cpp:69-79: the element sizes, returned as int
    int getDx() const { return (int)hx; }
    int getDy() const { return (int)hy; }
    int getDz() const { return (int)hz; }
    int getElementNumberFromPoints(const float &x, const float &y, const float &z) const
    {
        return phc_sem_element_from_point(ex, ey, ez, lx, ly, lz, x, y, z);
    }
    // node id of the grid point (ix, iy, iz) (SEMmesh.cpp:214-223 numbering)
    int node(int ix, int iy, int iz) const { return ix + iz * nx + iy * nx * nz; }
    // SEMmesh.cpp:203-236: globalNodesList(e, l), local node l = a + b (p+1) + c (p+1)^2 along x, z, y
    void globalNodesList(const int &numberOfElements, arrayInt &nodesList) const
    {
        const int q = order + 1, npe = getNumberOfPointsPerElement();
        nodesList = arrayInt(numberOfElements, npe);
        for (int e = 0; e < numberOfElements; ++e) {
            const int exi = e % ex, ezi = (e / ex) % ez, eyi = (ey == 0) ? 0 : e / (ex * ez);
            for (int l = 0; l < npe; ++l) {
                const int a = l % q, b = (l / q) % q, c = l / (q * q);
                nodesList(e, l) = node(exi * order + a, eyi * order + c, ezi * order + b);
            }
        }
    }
    // SEMmesh.cpp:240-262: interior nodes in node order
    int getListOfInteriorNodes(const int &numberOfInteriorNodes, vectorInt &list) const
    {
        list = vectorInt(numberOfInteriorNodes);
        int n = 0;
        for (int iy = (ny == 1 ? 0 : 1); iy < (ny == 1 ? 1 : ny - 1); ++iy)
            for (int iz = 1; iz < nz - 1; ++iz)
                for (int ix = 1; ix < nx - 1; ++ix) list[n++] = node(ix, iy, iz);
        return n;
    }
    // SEMmesh.cpp:288-312
    void getModel(const int &numberOfElements, vectorReal &m) const
    {
        m = vectorReal(numberOfElements);
        for (int e = 0; e < numberOfElements; ++e) m[e] = model[(size_t)e];
    }
    // SEMmesh.cpp:624-654: the node plane iy = ny/2 - 1 of level i1 as ASCII
    // "x z value" lines (x = ix * element size, z = iz * element size, as the
    // reference writes them), file "snapshot<step>"
    void saveSnapShot(const int indexTimeStep, const int i1, arrayReal const &u) const
    {
        const float dx = getDx(), dz = getDz();
        const int offset = (ny == 1) ? 0 : nx * nz * (ny / 2 - 1);
        std::ofstream snapFile("snapshot" + std::to_string(indexTimeStep), std::ios::out | std::ios::trunc);
        if (!snapFile) throw std::runtime_error("cannot open snapshot file");
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < nz; j++) snapFile << i * dx << " " << j * dz << " " << u(offset + i + j * nx, i1) << std::endl;
    }
};

class SEMsolver : public ColumnStore<float> {
  public:
    phc_sem *dev = nullptr;      // 3D mesh
    phc_sem2d *dev2 = nullptr;   // SEM2D mesh (ey = 0)
    int kernel = PHC_SEM_SEPARABLE;
    bool saveSnapShots = false;  // the reference's SEM_SAVE_SNAPSHOTS build option

    SEMsolver() = default;
    SEMsolver(const SEMsolver &) = delete;
    SEMsolver &operator=(const SEMsolver &) = delete;
    ~SEMsolver() override
    {
        if (dev) phc_sem_destroy(dev);
        if (dev2) phc_sem2d_destroy(dev2);
    }

    // SEMsolver.cpp:12-17: mesh, basis, geometry, mass -- built once in FP64
    // from the mesh's element speeds; the source element is myInfo's.
    void computeFEInit(SEMinfo &myInfo, const SEMmesh &mesh)
    {
        if (bound_) bound_->unbind();
        bound_ = nullptr;
        phc_sem_params p;
        std::memset(&p, 0, sizeof(p));
        p.ex = mesh.ex; p.ey = mesh.ey; p.ez = mesh.ez;
        p.lx = mesh.lx; p.ly = mesh.ly; p.lz = mesh.lz;
        p.order = myInfo.myOrderNumber;
        p.model = mesh.model.data();
        p.dt = myInfo.myTimeStep;
        p.rhs = nullptr;
        p.nsamples = 0;
        p.src_element = myInfo.myElementSource;
        p.kernel = kernel;
        p.weights_ref = 1;
        if (dev) phc_sem_destroy(dev);
        if (dev2) phc_sem2d_destroy(dev2);
        dev = nullptr;
        dev2 = nullptr;
        if (mesh.ey == 0) phc_check_sem(phc_sem2d_create(&p, &dev2), "phc_sem2d_create");
        else phc_check_sem(phc_sem_create(&p, &dev), "phc_sem_create");
        order_ = myInfo.myOrderNumber;
        nodes_ = mesh.getNumberOfNodes();
        ex_ = mesh.ex; ez_ = mesh.ez; ey_ = mesh.ey; nx_ = mesh.nx; nz_ = mesh.nz;
    }

    // SEMsolver.cpp:21-125
    void computeOneStep(const int &timeSample, const int &order, const int &nPointsPerElement, const int &i1,
                        const int &i2, SEMinfo &myInfo, const arrayReal &myRHSTerm, arrayReal const &myPnGlobal,
                        const vectorInt &myRhsElement)
    {
        if (!dev && !dev2) throw std::runtime_error("computeOneStep before computeFEInit");
        if (order != order_) throw std::runtime_error("computeOneStep: order differs from computeFEInit's");
        const int npe = (dev2 ? 1 : order + 1) * (order + 1) * (order + 1);
        if (nPointsPerElement != npe) throw std::runtime_error("computeOneStep: wrong nPointsPerElement");
        arrayReal &pn = bind(myPnGlobal);
        pn.push_host_changes();
        const int nr = myInfo.myNumberOfRHS;
        elems_.resize((size_t)nr);
        vals_.resize((size_t)nr);
        for (int r = 0; r < nr; ++r) {
            elems_[(size_t)r] = myRhsElement[r];
            vals_[(size_t)r] = myRHSTerm(r, timeSample);
        }
        phc_check_sem(dev2 ? phc_sem2d_compute_one_step_rhs(dev2, i1, i2, nr, elems_.data(), vals_.data())
                           : phc_sem_compute_one_step_rhs(dev, i1, i2, nr, elems_.data(), vals_.data()),
                      "computeOneStep");
        pn.device_changed(i1);
        pn.device_changed(i2);
    }

    // SEMsolver.cpp:127-142: every 100 steps, level i1 at the first node of
    // the source element
    void outputPnValues(const SEMmesh &mesh, const int &indexTimeStep, int &i1, int &myElementSource,
                        const arrayReal &pnGlobal)
    {
        if (indexTimeStep % 100 == 0) {
            const int e = myElementSource;
            const int node = (e % mesh.ex) * mesh.order + ((e / mesh.ex) % mesh.ez) * mesh.order * mesh.nx +
                             (mesh.ey == 0 ? 0 : (e / (mesh.ex * mesh.ez)) * mesh.order * mesh.nx * mesh.nz);
            printf("TimeStep=%d;  pnGlobal @ elementSource location %d after computeOneStep = %g\n", indexTimeStep,
                   myElementSource, pnGlobal(node, i1));
            if (saveSnapShots) mesh.saveSnapShot(indexTimeStep, i1, pnGlobal);
        }
    }

    // ---- additions (no reference counterpart): the fused, graph-captured
    // loop with device-side source and receiver traces
    void setSource(int element, const float *samples, int nsamples)
    {
        phc_check_sem(dev2 ? phc_sem2d_set_source(dev2, element, nsamples, samples)
                           : phc_sem_set_source(dev, element, nsamples, samples), "set_source");
    }
    void run(arrayReal const &pnGlobal, int it0, int n, int *a, int *b)
    {
        arrayReal &pn = bind(pnGlobal);
        pn.push_host_changes();
        phc_check_sem(dev2 ? phc_sem2d_run(dev2, it0, n, a, b) : phc_sem_run(dev, it0, n, a, b), "run");
        pn.device_changed(0);
        pn.device_changed(1);
    }
    void traces(float *src, float *rec)
    {
        phc_check_sem(dev2 ? phc_sem2d_get_traces(dev2, src, rec) : phc_sem_get_traces(dev, src, rec), "traces");
    }
    void setReceivers(int n, const int *nodes)
    {
        phc_check_sem(dev2 ? phc_sem2d_set_receivers(dev2, n, nodes) : phc_sem_set_receivers(dev, n, nodes),
                      "receivers");
    }
    void synchronize() { phc_check_sem(dev2 ? phc_sem2d_synchronize(dev2) : phc_sem_synchronize(dev), "synchronize"); }

    // ColumnStore: the levels of the bound pnGlobal
    void pull(int col, float *host) override
    {
        phc_check_sem(dev2 ? phc_sem2d_get_field(dev2, col, host) : phc_sem_get_field(dev, col, host), "get_field");
    }
    void push(int col, const float *host) override
    {
        phc_check_sem(dev2 ? phc_sem2d_set_field(dev2, col, host) : phc_sem_set_field(dev, col, host), "set_field");
    }

  private:
    // The caller's pnGlobal becomes the device's view at the first step; the
    // reference modifies it through the const reference as well
    // (SEMsolver.cpp:30-33, CREATEVIEWS).
    arrayReal &bind(arrayReal const &pnGlobal)
    {
        arrayReal &pn = const_cast<arrayReal &>(pnGlobal);
        if (bound_ == &pn) return pn;
        if (bound_) throw std::runtime_error("computeOneStep: a different pnGlobal than the earlier steps'");
        if (pn.rows() != nodes_ || pn.cols() != 2)
            throw std::runtime_error("pnGlobal is not the (numberOfNodes, 2) array");
        pn.bind(this);
        bound_ = &pn;
        return pn;
    }
    arrayReal *bound_ = nullptr;
    int order_ = 0, nodes_ = 0, ex_ = 0, ey_ = 0, ez_ = 0, nx_ = 0, nz_ = 0;
    std::vector<int> elems_;
    std::vector<float> vals_;
};
