// FDTD.hpp -- C++ drop-in for the reference FD proxy's classes, backed by
// libphc.so (the two wavefield levels live on the GPU).
//
//   FDTDGRIDS / FDTDMODELS     src/fd/utilsFDTD/FDTDdata.hpp:6-42
//   FDTDInit                   src/fd/utilsFDTD/FDTDinit.hpp:9-397
//   FDTDKernel                 src/fd/utilsFDTD/FDTDkernels.hpp:4-131
//   FDTDUtils                  src/fd/utilsFDTD/FDTDutils.hpp:13-76, 193-232
//
// Same class names, members, method signatures and argv flags as the
// reference, so its driver (FDTDmain.cpp:22-81) compiles and runs unchanged:
//   init_geometry(argc, argv, grids); init_coefficients(grids, models);
//   init_source(models); init_models(grids, models);
//   defineSpongeBoundary(grids, models.spongeArray);
//   addRHS(grids, it, models, i2, models.pnGlobal);
//   computeOneStepSB / computeOneStepPML(grids, models, i1, i2);
//   output(grids, models.pnGlobal, it, i2, saveSnapShots).
// The device state is created at the first step call, from what these
// calls left in FDTDMODELS.  FDTDMODELS::pnGlobal is the reference's
// (IDX3_l node, level) array; its levels live on the GPU and the host copy
// follows the ColumnStore rules of phc_arrays.hpp (read after a step:
// refreshed; written by the caller: sent back before the next step).
// addRHS injects RHSTerm[itSample] as the caller holds it at that call, at
// the grid's source point as it is at that call.
//
// Additional flags (not in the reference): -f0 F, -cfl C, -timemax T,
// -sponge W, -layers N SEED (seeded layered model), -homogeneous (vmin
// everywhere, the model of the published test_20240512 runs), -fp32
// (all-FP32 arithmetic), -fused (graph-captured 50-step runs), -nsteps N.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "phc.h"
#include "phc_arrays.hpp"
#include "phc_setup.h"

inline void phc_check(phc_status s, const char *what)
{
    if (s != PHC_OK) throw std::runtime_error(std::string(what) + ": " + phc_last_error());
}

struct FDTDGRIDS {
    int nx, ny, nz;
    int xs, ys, zs;
    int lx, ly, lz;
    float dx, dy, dz;
    int ntaperx, ntapery, ntaperz;
    int ndampx, ndampy, ndampz;
    float hdx_2, hdy_2, hdz_2;
    int x1, x2, x3, x4, x5, x6;
    int y1, y2, y3, y4, y5, y6;
    int z1, z2, z3, z4, z5, z6;
};

// The reference's spongeArray (FDTDinit.hpp:319-396) holds nx*ny*nz values
// that follow a face-slab rule; here the rule itself: width W and alpha,
// evaluated on the device and, for host reads, by operator().  Undefined
// until defineSpongeBoundary.
struct SpongeArray {
    bool defined = false;
    int width = 0;
    float alpha = 0.f;
    int nx = 0, ny = 0, nz = 0;
    // value at IDX3(i,j,k) = NZ*NY*i + NZ*j + k: exp(alpha (W-d)^2) on the
    // face slabs (the other two coordinates in [W, n-W)), 1 elsewhere
    float operator()(int64_t idx3) const
    {
        if (!defined) throw std::runtime_error("spongeArray read before defineSpongeBoundary");
        const int k = (int)(idx3 % nz), j = (int)((idx3 / nz) % ny), i = (int)(idx3 / ((int64_t)nz * ny));
        const int W = width;
        const bool inI = i >= W && i < nx - W, inJ = j >= W && j < ny - W, inK = k >= W && k < nz - W;
        auto f = [&](int a, int n) {
            const int d = (a >= n - W) ? W - (n - a) : ((a < W) ? W - a : -1);
            return d >= 0 ? (float)std::exp((double)alpha * std::pow((double)d, 2.0)) : 1.f;
        };
        float s = 1.f;
        if (inJ && inK && !inI) s = f(i, nx);
        if (inK && inI && !inJ) s = f(j, ny);
        if (inJ && inI && !inK) s = f(k, nz);
        return s;
    }
    float operator[](int64_t idx3) const { return (*this)(idx3); }
};

// libphc.so handle behind FDTDMODELS, and the pnGlobal levels it holds.
struct FdDevice : ColumnStore<float> {
    phc_fd *h = nullptr;
    int nx = 0, ny = 0, nz = 0, lx = 0, ly = 0, lz = 0;
    FdDevice() = default;
    FdDevice(const FdDevice &) = delete;
    FdDevice &operator=(const FdDevice &) = delete;
    ~FdDevice() override
    {
        if (h) phc_fd_destroy(h);
    }
    void pull(int col, float *host) override { phc_check(phc_fd_get_field(h, col, host), "get_field"); }
    void push(int col, const float *host) override { phc_check(phc_fd_set_field(h, col, host), "set_field"); }
    bool read_one(int64_t row, int col, float *v) override
    {
        // row = IDX3_l(i,j,k) (FDTDmacros.hpp:14)
        const int64_t NZe = nz + 2 * lz, NYe = ny + 2 * ly;
        const int k = (int)(row % NZe) - lz, j = (int)((row / NZe) % NYe) - ly, i = (int)(row / (NZe * NYe)) - lx;
        phc_check(phc_fd_read_point(h, col, i, j, k, v), "read_point");
        return true;
    }
};

struct FDTDMODELS {
    double coef0 = 0.0;
    vectorReal coefx, coefy, coefz, RHSTerm;
    SpongeArray spongeArray;
    vectorReal vp;
    // The reference also allocates phi, eta, pnp1 and pn (FDTDinit.hpp:255-267).
    // Here phi lives on the device and eta is three 1D profiles (etaProfile);
    // these stay empty.
    vectorReal phi, eta, pnp1, pn;
    arrayReal pnGlobal;

    // ---- device configuration recorded by FDTDInit::init_models
    int arith = PHC_FD_ARITH_REF;
    bool usePML = false;
    int ndamp[3] = {0, 0, 0};
    std::vector<float> etaProfile[3];
    float hd2[3] = {0.f, 0.f, 0.f};
    std::unique_ptr<FdDevice> dev;

    // The device state, created at the first step call (levels at
    // pnGlobal's values, the sponge as defineSpongeBoundary defined it, the
    // PML when init_models set one up).
    phc_fd *device(const FDTDGRIDS &g)
    {
        if (dev) return dev->h;
        phc_fd_params p;
        std::memset(&p, 0, sizeof(p));
        p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
        p.lx = g.lx; p.ly = g.ly; p.lz = g.lz;
        p.xs = g.xs; p.ys = g.ys; p.zs = g.zs;
        if (coefx.size() < g.lx + 1 || coefy.size() < g.ly + 1 || coefz.size() < g.lz + 1)
            throw std::runtime_error("coefficients missing: call init_coefficients first");
        if (vp.size() != g.nx * g.ny * g.nz) throw std::runtime_error("vp missing: call init_models first");
        const int64_t ext = (int64_t)(g.nx + 2 * g.lx) * (g.ny + 2 * g.ly) * (g.nz + 2 * g.lz);
        if (pnGlobal.rows() != ext || pnGlobal.cols() != 2)
            throw std::runtime_error("pnGlobal is not the (IDX3_l, 2) array init_models allocates");
        p.coefx = coefx.data(); p.coefy = coefy.data(); p.coefz = coefz.data();
        p.coef0 = coef0;
        p.vp = vp.data();
        p.rhs = RHSTerm.data();
        p.nsamples = RHSTerm.size();
        if (usePML) {
            p.sponge_width = 0;  // computeOneStepPML applies no sponge
        } else {
            if (!spongeArray.defined) throw std::runtime_error("sponge undefined: call defineSpongeBoundary first");
            p.sponge_width = spongeArray.width;
            p.sponge_alpha = spongeArray.alpha;
        }
        p.init_value = pnGlobal.fill_value();
        p.arith = arith;
        auto d = std::make_unique<FdDevice>();
        d->nx = g.nx; d->ny = g.ny; d->nz = g.nz; d->lx = g.lx; d->ly = g.ly; d->lz = g.lz;
        phc_check(phc_fd_create(&p, &d->h), "phc_fd_create");
        if (usePML)
            phc_check(phc_fd_set_pml(d->h, ndamp, etaProfile[0].data(), etaProfile[1].data(),
                                     etaProfile[2].data(), hd2), "set_pml");
        dev = std::move(d);
        pnGlobal.bind(dev.get());
        return dev->h;
    }
    void check_field(const arrayReal &pn) const
    {
        if (&pn != &pnGlobal)
            throw std::runtime_error("the wavefield passed is not FDTDMODELS::pnGlobal, whose levels the device holds");
    }
};

struct FDTDUtils {
    // FDTDutils.hpp:13-58
    void init_coef(int L, float dx, vectorReal &coef) { phc_check(phc_fd_init_coef(L, dx, coef.data()), "init_coef"); }
    // FDTDutils.hpp:59-76 (cfl 0.8 in the reference)
    float compute_dt_sch(float vmax, vectorReal const &cx, vectorReal const &cy, vectorReal const &cz, float cfl = 0.8f)
    {
        return phc_fd_compute_dt(vmax, cx.data(), cx.size() - 1, cy.data(), cy.size() - 1, cz.data(), cz.size() - 1,
                                 cfl);
    }
    // FDTDutils.hpp:193-210: every 50 steps print the level at the source
    // point, optionally write the x-z snapshot
    void output(FDTDGRIDS &g, arrayReal const &pnGlobal, int itSample, const int &i1, const bool saveSnapShots)
    {
        if (itSample % 50 == 0) {
            const int64_t src = (int64_t)(g.nz + 2 * g.lz) * (g.ny + 2 * g.ly) * (g.xs + g.lx) +
                                (int64_t)(g.nz + 2 * g.lz) * (g.ys + g.ly) + (g.zs + g.lz);
            printf("TimeStep=%d\t; Pressure value at source [%d %d %d] = %f\n", itSample, g.xs, g.ys, g.zs,
                   pnGlobal.value(src, i1));
            if (saveSnapShots) write_io(g, 0, g.nx, g.ny / 2, g.ny / 2, 0, g.nz, pnGlobal, itSample, i1);
        }
    }
    // FDTDutils.hpp:212-232: raw FP32 values for k in [z0,z1), j in [y0,y1],
    // i in [x0,x1), i fastest, file snapshot_it_<istep>.H@
    void write_io(FDTDGRIDS &g, int x0, int x1, int y0, int y1, int z0, int z1, arrayReal const &pnGlobal,
                  int istep, const int &i1)
    {
        char name[64];
        snprintf(name, sizeof(name), "snapshot_it_%d.H@", istep);
        FILE *f = fopen(name, "wb");
        if (!f) throw std::runtime_error("cannot open snapshot file");
        // pnGlobal(IDX3_l(i,j,k), i1): the first read refreshes the host copy of level i1
        const int64_t NZe = g.nz + 2 * g.lz, NYe = g.ny + 2 * g.ly;
        for (int k = z0; k < z1; ++k)
            for (int j = y0; j < y1 + 1; ++j)
                for (int i = x0; i < x1; ++i) {
                    const float v = pnGlobal(NZe * NYe * (i + g.lx) + NZe * (j + g.ly) + (k + g.lz), i1);
                    fwrite(&v, sizeof(float), 1, f);
                }
        fclose(f);
    }
};

struct FDTDKernel {
    // FDTDkernels.hpp:121-128: pnGlobal(src, cb) += vp[src] * RHSTerm[itSample]
    int addRHS(FDTDGRIDS &g, const int itSample, FDTDMODELS &m, int &cb, arrayReal const &pnGlobal)
    {
        m.check_field(pnGlobal);
        if (itSample < 0 || itSample >= m.RHSTerm.size())
            throw std::out_of_range("addRHS: time sample outside RHSTerm");
        phc_fd *h = m.device(g);
        m.pnGlobal.push_host_changes();
        phc_check(phc_fd_add_rhs_value(h, cb, g.xs, g.ys, g.zs, m.RHSTerm[itSample]), "addRHS");
        m.pnGlobal.device_changed(cb);
        return 0;
    }
    // FDTDkernels.hpp:75-84 (inner3D + applySponge in one launch)
    int computeOneStepSB(FDTDGRIDS &g, FDTDMODELS &m, int &ca, int &cb)
    {
        phc_fd *h = m.device(g);
        m.pnGlobal.push_host_changes();
        phc_check(phc_fd_compute_one_step_sb(h, ca, cb), "computeOneStepSB");
        m.pnGlobal.device_changed(ca);
        m.pnGlobal.device_changed(cb);
        return 0;
    }
    // FDTDkernels.hpp:86-104 (pml3D on the slabs, inner3D inside)
    int computeOneStepPML(FDTDGRIDS &g, FDTDMODELS &m, int &ca, int &cb)
    {
        phc_fd *h = m.device(g);
        m.pnGlobal.push_host_changes();
        phc_check(phc_fd_compute_one_step_pml(h, ca, cb), "computeOneStepPML");
        m.pnGlobal.device_changed(ca);
        return 0;
    }
};

struct FDTDInit {
    FDTDUtils myFDTDUtils;
    int sourceOrder = 1;
    int ncoefsX = 5, ncoefsY = 5, ncoefsZ = 5;
    float f0 = 15.f;
    float fmax = 2.5f * 15.f;
    float timeMax = 1.0f;
    float timeStep = 0.f;
    int nSamples = 0;
    float vmin = 1500.f, vmax = 4500.f;
    int i1 = 0, i2 = 1;
    bool saveSnapShots = false, usePML = false;
    vectorReal RHSTerm;
    // additional options
    float cfl = 0.8f;
    int spongeWidth = 20;
    float spongeAlpha = -0.00015f;
    int layers = 0;
    unsigned long long layerSeed = 0;
    int arith = PHC_FD_ARITH_REF;
    bool fused = false;
    bool homogeneous = false;
    int nsteps = -1;

    // FDTDinit.hpp:36-200
    void init_geometry(int argc, char *argv[], FDTDGRIDS &g)
    {
        g.nx = 150; g.ny = g.nx; g.nz = g.nx;
        g.xs = g.nx / 2; g.ys = g.ny / 2; g.zs = g.nz / 2;
        g.lx = g.ly = g.lz = 4;
        g.dx = g.dy = g.dz = 10;
        if (argc > 1) {
            int i = 1;
            while (i < argc) {
                const std::string a = argv[i];
                auto need = [&](int n) {
                    if (i + n >= argc) throw std::runtime_error("missing value after " + a);
                };
                if (a == "-nx") { need(1); g.nx = g.ny = g.nz = atoi(argv[i + 1]); i += 2; }
                else if (a == "-ny") { need(1); g.ny = atoi(argv[i + 1]); i += 2; }
                else if (a == "-nz") { need(1); g.nz = atoi(argv[i + 1]); i += 2; }
                else if (a == "-dx") { need(1); g.dx = g.dy = g.dz = atoi(argv[i + 1]); i += 2; }
                else if (a == "-dy") { need(1); g.dy = atoi(argv[i + 1]); i += 2; }
                else if (a == "-dz") { need(1); g.dz = atoi(argv[i + 1]); i += 2; }
                else if (a == "-lx") { need(1); g.lx = g.ly = g.lz = atoi(argv[i + 1]); i += 2; }
                else if (a == "-ly") { need(1); g.ly = atoi(argv[i + 1]); i += 2; }
                else if (a == "-lz") { need(1); g.lz = atoi(argv[i + 1]); i += 2; }
                else if (a == "-xs") { need(1); g.xs = g.ys = g.zs = atoi(argv[i + 1]); i += 2; }
                else if (a == "-ys") { need(1); g.ys = atoi(argv[i + 1]); i += 2; }
                else if (a == "-zs") { need(1); g.zs = atoi(argv[i + 1]); i += 2; }
                else if (a == "-savesnapshots") { saveSnapShots = true; i += 1; }
                else if (a == "-usePML") { usePML = true; i += 1; }
                else if (a == "-f0") { need(1); f0 = (float)atof(argv[i + 1]); i += 2; }
                else if (a == "-cfl") { need(1); cfl = (float)atof(argv[i + 1]); i += 2; }
                else if (a == "-timemax") { need(1); timeMax = (float)atof(argv[i + 1]); i += 2; }
                else if (a == "-sponge") { need(1); spongeWidth = atoi(argv[i + 1]); i += 2; }
                else if (a == "-layers") { need(2); layers = atoi(argv[i + 1]); layerSeed = strtoull(argv[i + 2], nullptr, 10); i += 3; }
                else if (a == "-fp32") { arith = PHC_FD_ARITH_FP32; i += 1; }
                else if (a == "-fused") { fused = true; i += 1; }
                else if (a == "-homogeneous") { homogeneous = true; i += 1; }
                else if (a == "-nsteps") { need(1); nsteps = atoi(argv[i + 1]); i += 2; }
                else throw std::runtime_error("unknown argument " + a);
            }
        } else {
            printf("usage fd_exe -nx XXX -ny YYY -nz ZZZZ -lx LX -ly LY -lz LZ -dx DX -dy DY -dz DZ  ");
            printf("-xs XS -ys YS -zs ZS -savesnapshots (false) -usePML (false) \n ");
            printf("if no argument running with default values\n");
        }
        for (int l : {g.lx, g.ly, g.lz})
            if (l < 1 || l > 6) throw std::runtime_error("half stencil length outside 1..6");
        ncoefsX = g.lx + 1; ncoefsY = g.ly + 1; ncoefsZ = g.lz + 1;
        fmax = 2.5f * f0;
        printf("Number of grids: nx=%d, ny=%d, nz=%d\n", g.nx, g.ny, g.nz);
        printf("Source location: xs=%d, ys=%d, zs=%d\n", g.xs, g.ys, g.zs);
        g.ntaperx = g.ntapery = g.ntaperz = 4;
        g.hdx_2 = 1. / (4. * g.dx * g.dx);
        g.hdy_2 = 1. / (4. * g.dy * g.dy);
        g.hdz_2 = 1. / (4. * g.dz * g.dz);
        printf("usePML=%d\n", usePML);
        if (usePML) {  // FDTDinit.hpp:167-172: ntaper * lambdamax / d, lambdamax = vmin / fmax
            g.ndampx = phc_fd_pml_width(vmin, f0, g.dx);
            g.ndampy = phc_fd_pml_width(vmin, f0, g.dy);
            g.ndampz = phc_fd_pml_width(vmin, f0, g.dz);
        } else {
            g.ndampx = g.ndampy = g.ndampz = 0;
        }
        g.x1 = 0; g.x2 = g.ndampx; g.x3 = g.ndampx; g.x4 = g.nx - g.ndampx; g.x5 = g.nx - g.ndampx; g.x6 = g.nx;
        g.y1 = 0; g.y2 = g.ndampy; g.y3 = g.ndampy; g.y4 = g.ny - g.ndampy; g.y5 = g.ny - g.ndampy; g.y6 = g.ny;
        g.z1 = 0; g.z2 = g.ndampz; g.z3 = g.ndampz; g.z4 = g.nz - g.ndampz; g.z5 = g.nz - g.ndampz; g.z6 = g.nz;
    }

    // FDTDinit.hpp:202-233
    void init_coefficients(FDTDGRIDS &g, FDTDMODELS &m)
    {
        m.coefx = allocateVector<vectorReal>(ncoefsX, "coefx");
        m.coefy = allocateVector<vectorReal>(ncoefsY, "coefy");
        m.coefz = allocateVector<vectorReal>(ncoefsZ, "coefz");
        myFDTDUtils.init_coef(g.lx, g.dx, m.coefx);
        myFDTDUtils.init_coef(g.ly, g.dy, m.coefy);
        myFDTDUtils.init_coef(g.lz, g.dz, m.coefz);
        m.coef0 = phc_fd_coef0(m.coefx.data(), g.lx, m.coefy.data(), g.ly, m.coefz.data(), g.lz);
        timeStep = myFDTDUtils.compute_dt_sch(vmax, m.coefx, m.coefy, m.coefz, cfl);
        nSamples = (int)(timeMax / timeStep);
        if (nsteps >= 0 && nsteps < nSamples) nSamples = nsteps;
        printf("init coefs done\n");
    }

    // FDTDinit.hpp:235-247
    void init_source(FDTDMODELS &m)
    {
        m.RHSTerm = allocateVector<vectorReal>(nSamples, "RHSTerm");
        phc_check(phc_source_term(nSamples, timeStep, f0, sourceOrder, PHC_EXP_DOUBLE, m.RHSTerm.data()),
                  "source term");
    }

    // FDTDinit.hpp:249-317: model, levels at 1e-6 (halo included), PML eta
    void init_models(FDTDGRIDS &g, FDTDMODELS &m)
    {
        m.vp = allocateVector<vectorReal>(g.nx * g.ny * g.nz, "vp");
        if (homogeneous) {
            const float v0 = ((vmin * vmin) * timeStep) * timeStep;
            for (int n = 0; n < m.vp.size(); ++n) m.vp[n] = v0;
        } else if (layers > 0)
            phc_check(phc_fd_model_layered(g.nx, g.ny, g.nz, layers, vmin, vmax, layerSeed, timeStep,
                                           nullptr, m.vp.data()), "layered model");
        else
            phc_check(phc_fd_model_two_layer(g.nx, g.ny, g.nz, vmin, timeStep, m.vp.data()), "model");
        const int ext = (g.nx + 2 * g.lx) * (g.ny + 2 * g.ly) * (g.nz + 2 * g.lz);
        m.pnGlobal = allocateArray2D<arrayReal>(ext, 2, "pnGlobal");
        m.pnGlobal.fill(0.000001f);
        m.arith = arith;
        m.usePML = usePML;
        if (usePML) {
            // FDTDutils.hpp:129-171 init_eta: per-axis quadratic profiles whose
            // sum is the reference's eta volume
            const int nd[3] = {g.ndampx, g.ndampy, g.ndampz};
            const int n[3] = {g.nx, g.ny, g.nz};
            const float d[3] = {g.dx, g.dy, g.dz};
            for (int a = 0; a < 3; ++a) {
                m.ndamp[a] = nd[a];
                m.etaProfile[a].assign(n[a] + 2, 0.f);
                if (nd[a] > 0)
                    phc_check(phc_fd_pml_profile(n[a], nd[a], phc_fd_pml_scale(timeStep, vmax, nd[a], d[a]),
                                                 m.etaProfile[a].data()), "pml profile");
            }
            m.hd2[0] = g.hdx_2; m.hd2[1] = g.hdy_2; m.hd2[2] = g.hdz_2;
        }
    }

    // FDTDinit.hpp:319-396 (width 20 and alpha -1.5e-4 in the reference;
    // -sponge W sets the width here)
    void defineSpongeBoundary(FDTDGRIDS &g, SpongeArray &spongeArray)
    {
        spongeArray.defined = true;
        spongeArray.width = spongeWidth;
        spongeArray.alpha = spongeAlpha;
        spongeArray.nx = g.nx; spongeArray.ny = g.ny; spongeArray.nz = g.nz;
    }
};
