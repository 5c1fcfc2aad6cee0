// dropin_semantics.cpp -- checks that the drop-in classes honour what the
// reference's callers hand over at each call: host edits of pnGlobal between
// steps, the RHS values as they stand at each step, and the source point of
// FDTDGRIDS.  Each case runs twice on the same inputs: once through the
// drop-in classes with the reference's calls (FDTDmain.cpp:55-81,
// SEMproxy.cpp:45-55), once through the C ABI with the same edits applied by
// hand.  The two must agree bit for bit.  Exit status 0 when they do.
#include <cstdio>
#include <cstring>
#include <vector>

#include "FDTD.hpp"
#include "SEM.hpp"

static int fd_case()
{
    const char *argv[] = {"fd", "-nx", "32", "-ny", "28", "-nz", "40", "-lx", "4", "-ly", "3", "-lz", "2",
                          "-xs", "16", "-ys", "14", "-zs", "20", "-sponge", "6", "-f0", "30", "-nsteps", "40"};
    const int argc = sizeof(argv) / sizeof(argv[0]);
    FDTDGRIDS g;
    FDTDMODELS m;
    FDTDInit init;
    FDTDKernel kernel;
    FDTDUtils utils;
    init.init_geometry(argc, const_cast<char **>(argv), g);
    init.init_coefficients(g, m);
    init.init_source(m);
    init.init_models(g, m);
    init.defineSpongeBoundary(g, m.spongeArray);

    // the same problem through the C ABI
    std::vector<float> rhs(m.RHSTerm.data(), m.RHSTerm.data() + m.RHSTerm.size());
    phc_fd_params p;
    std::memset(&p, 0, sizeof(p));
    p.nx = g.nx; p.ny = g.ny; p.nz = g.nz; p.lx = g.lx; p.ly = g.ly; p.lz = g.lz;
    p.xs = g.xs; p.ys = g.ys; p.zs = g.zs;
    p.coefx = m.coefx.data(); p.coefy = m.coefy.data(); p.coefz = m.coefz.data(); p.coef0 = m.coef0;
    p.vp = m.vp.data(); p.rhs = rhs.data(); p.nsamples = (int)rhs.size();
    p.sponge_width = 6; p.sponge_alpha = -0.00015f; p.init_value = 0.000001f; p.arith = PHC_FD_ARITH_REF;
    phc_fd *b = nullptr;
    phc_check(phc_fd_create(&p, &b), "create");
    const int64_t NZe = g.nz + 2 * g.lz, NYe = g.ny + 2 * g.ly, NXe = g.nx + 2 * g.lx;
    const int64_t ext = NXe * NYe * NZe;
    auto idx = [&](int i, int j, int k) { return NZe * NYe * (i + g.lx) + NZe * (j + g.ly) + (k + g.lz); };
    int bi1 = 0, bi2 = 1;
    std::vector<float> f((size_t)ext);
    for (int it = 0; it < init.nSamples; ++it) {
        if (it == 20) {
            // host edits between steps: a point of the current level, a later
            // source sample, and the source point itself
            m.pnGlobal(idx(10, 11, 12), init.i2) += 0.5f;
            m.RHSTerm[25] = 0.3f;
            g.xs = 12;
            phc_check(phc_fd_get_field(b, bi2, f.data()), "get");
            f[(size_t)idx(10, 11, 12)] += 0.5f;
            phc_check(phc_fd_set_field(b, bi2, f.data()), "set");
            rhs[25] = 0.3f;
        }
        kernel.addRHS(g, it, m, init.i2, m.pnGlobal);
        kernel.computeOneStepSB(g, m, init.i1, init.i2);
        std::swap(init.i1, init.i2);
        utils.output(g, m.pnGlobal, it, init.i2, false);

        phc_check(phc_fd_add_rhs_value(b, bi2, it < 20 ? 16 : 12, 14, 20, rhs[(size_t)it]), "add");
        phc_check(phc_fd_compute_one_step_sb(b, bi1, bi2), "step");
        std::swap(bi1, bi2);
    }
    int bad = 0;
    for (int l = 0; l < 2; ++l) {
        const int la = l == 0 ? init.i2 : init.i1, lb = l == 0 ? bi2 : bi1;
        phc_check(phc_fd_get_field(b, lb, f.data()), "get");
        const arrayReal &pn = m.pnGlobal;
        for (int64_t n = 0; n < ext; ++n)
            if (pn(n, la) != f[(size_t)n]) {
                if (bad < 5) printf("fd mismatch level %d at %lld: %.9g vs %.9g\n", l, (long long)n, pn(n, la), f[(size_t)n]);
                ++bad;
            }
    }
    phc_fd_destroy(b);
    printf("fd case: %d mismatches\n", bad);
    return bad;
}

static int sem_case(bool dim2, int kernel_kind)
{
    SEMinfo info;
    info.myOrderNumber = 3;
    info.myTimeStep = 0.001f;
    info.f0 = 20.f;
    info.myNumSamples = 40;
    info.myNumberOfRHS = 2;
    const int ex = 6;
    SEMmesh mesh(ex, ex, ex, 600.f, 600.f, 600.f, info.myOrderNumber, dim2);
    for (size_t e = 0; e < mesh.model.size(); ++e) mesh.model[e] = 1500.f + 10.f * (float)(e % 7);
    const int q = info.myOrderNumber + 1;
    info.nPointsPerElement = dim2 ? q * q : q * q * q;
    info.numberOfNodes = mesh.getNumberOfNodes();
    info.numberOfElements = mesh.getNumberOfElements();
    arrayReal RHS = allocateArray2D<arrayReal>(info.myNumberOfRHS, info.myNumSamples, "RHSTerm");
    vectorInt rhsElement = allocateVector<vectorInt>(info.myNumberOfRHS, "rhsElement");
    arrayReal pnGlobal = allocateArray2D<arrayReal>(info.numberOfNodes, 2, "pnGlobal");
    SolverUtils u;
    std::vector<float> st = u.computeSourceTerm(info.myNumSamples, info.myTimeStep, info.f0, info.sourceOrder);
    for (int j = 0; j < info.myNumSamples; ++j) {
        RHS(0, j) = st[(size_t)j];
        RHS(1, j) = -0.5f * st[(size_t)j];
    }
    rhsElement[0] = mesh.getElementNumberFromPoints(301.f, 301.f, 301.f);
    rhsElement[1] = mesh.getElementNumberFromPoints(150.f, 301.f, 450.f);
    info.myElementSource = rhsElement[0];
    SEMsolver solver;
    solver.kernel = kernel_kind;
    solver.computeFEInit(info, mesh);

    // the same problem through the C ABI
    phc_sem_params p;
    std::memset(&p, 0, sizeof(p));
    p.ex = mesh.ex; p.ey = mesh.ey; p.ez = mesh.ez; p.lx = mesh.lx; p.ly = mesh.ly; p.lz = mesh.lz;
    p.order = info.myOrderNumber; p.model = mesh.model.data(); p.dt = info.myTimeStep;
    p.src_element = rhsElement[0]; p.kernel = kernel_kind; p.weights_ref = 1;
    phc_sem *b = nullptr;
    phc_sem2d *b2 = nullptr;
    if (dim2) phc_check_sem(phc_sem2d_create(&p, &b2), "create2d");
    else phc_check_sem(phc_sem_create(&p, &b), "create");
    auto get = [&](int l, float *h) { phc_check_sem(dim2 ? phc_sem2d_get_field(b2, l, h) : phc_sem_get_field(b, l, h), "get"); };
    auto set = [&](int l, const float *h) { phc_check_sem(dim2 ? phc_sem2d_set_field(b2, l, h) : phc_sem_set_field(b, l, h), "set"); };
    std::vector<float> f((size_t)info.numberOfNodes);
    std::vector<float> vals(2);
    std::vector<int> elems = {rhsElement[0], rhsElement[1]};
    const int edit_node = mesh.node(7, dim2 ? 0 : 8, 9);
    int i1 = 0, i2 = 1, bi1 = 0, bi2 = 1;
    for (int it = 0; it < info.myNumSamples; ++it) {
        if (it == 15) {
            pnGlobal(edit_node, i2) += 0.25f;
            RHS(1, 22) = 0.7f;
            get(bi2, f.data());
            f[(size_t)edit_node] += 0.25f;
            set(bi2, f.data());
        }
        solver.computeOneStep(it, info.myOrderNumber, info.nPointsPerElement, i1, i2, info, RHS, pnGlobal, rhsElement);
        solver.outputPnValues(mesh, it, i1, info.myElementSource, pnGlobal);
        std::swap(i1, i2);
        vals[0] = st[(size_t)it];
        vals[1] = (it == 22) ? 0.7f : -0.5f * st[(size_t)it];
        phc_check_sem(dim2 ? phc_sem2d_compute_one_step_rhs(b2, bi1, bi2, 2, elems.data(), vals.data())
                           : phc_sem_compute_one_step_rhs(b, bi1, bi2, 2, elems.data(), vals.data()), "step");
        std::swap(bi1, bi2);
    }
    int bad = 0;
    for (int l = 0; l < 2; ++l) {
        const int la = l == 0 ? i2 : i1, lb = l == 0 ? bi2 : bi1;
        get(lb, f.data());
        const arrayReal &pn = pnGlobal;
        for (int n = 0; n < info.numberOfNodes; ++n)
            if (pn(n, la) != f[(size_t)n]) {
                if (bad < 5) printf("sem mismatch level %d at %d: %.9g vs %.9g\n", l, n, pn(n, la), f[(size_t)n]);
                ++bad;
            }
    }
    if (b) phc_sem_destroy(b);
    if (b2) phc_sem2d_destroy(b2);
    printf("sem case (2d %d, kernel %d): %d mismatches\n", (int)dim2, kernel_kind, bad);
    return bad;
}

int main()
{
    try {
        int bad = fd_case();
        bad += sem_case(false, PHC_SEM_SEPARABLE);
        bad += sem_case(false, PHC_SEM_ELEMENT_COLOR);
        bad += sem_case(true, PHC_SEM_SEPARABLE);
        printf(bad ? "DROPIN MISMATCH\n" : "DROPIN OK\n");
        return bad ? 1 : 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "dropin_semantics: %s\n", e.what());
        return 2;
    }
}
