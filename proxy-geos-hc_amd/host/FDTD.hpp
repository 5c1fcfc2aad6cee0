// FDTD.hpp -- C++ drop-in for the reference FD proxy's classes, backed by
// libphc.so (the fields live on the GPU; host copies are made only for output).
//
//   FDTDGRIDS / FDTDMODELS   src/fd/utilsFDTD/FDTDdata.hpp:6-42
//   FDTDInit                 src/fd/utilsFDTD/FDTDinit.hpp:9-397
//   FDTDKernel               src/fd/utilsFDTD/FDTDkernels.hpp:4-131
//   FDTDUtils::output/write_io src/fd/utilsFDTD/FDTDutils.hpp:193-232
//
// Same class names, method names, argument meaning and argv flags (including
// -usePML: the PML of FDTDKernel::computeOneStepPML).  Additional
// flags (not in the reference): -f0 F, -cfl C, -timemax T, -sponge W,
// -layers N SEED (seeded layered model), -homogeneous (vmin everywhere, the
// model of the published test_20240512 runs), -fp32 (all-FP32 arithmetic),
// -fused (graph-captured 50-step runs), -nsteps N.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "phc.h"
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

struct FDTDMODELS {
    double coef0 = 0.0;
    std::vector<float> coefx, coefy, coefz, RHSTerm, vp;
    // the two wavefield levels pnGlobal(:,0), pnGlobal(:,1) live in `dev`
    phc_fd *dev = nullptr;
    ~FDTDMODELS()
    {
        if (dev) phc_fd_destroy(dev);
    }
};

struct FDTDUtils {
    // FDTDutils.hpp:193-210: every 50 steps print the current level at the
    // source point, optionally write the x-z snapshot
    void output(FDTDGRIDS &g, FDTDMODELS &m, int itSample, const int &i1, bool saveSnapShots)
    {
        if (itSample % 50 == 0) {
            float v = 0.f;
            phc_check(phc_fd_read_point(m.dev, i1, g.xs, g.ys, g.zs, &v), "read_point");
            printf("TimeStep=%d\t; Pressure value at source [%d %d %d] = %f\n", itSample, g.xs, g.ys,
                   g.zs, v);
            if (saveSnapShots) write_io(g, m, itSample, i1);
        }
    }
    // FDTDutils.hpp:212-232: raw FP32, x fastest, the plane j = ny/2
    void write_io(FDTDGRIDS &g, FDTDMODELS &m, int istep, const int &i1)
    {
        std::vector<float> slice((size_t)g.nx * g.nz);
        phc_check(phc_fd_get_slice_xz(m.dev, i1, g.ny / 2, slice.data()), "get_slice_xz");
        char name[64];
        snprintf(name, sizeof(name), "snapshot_it_%d.H@", istep);
        FILE *f = fopen(name, "wb");
        if (!f) throw std::runtime_error("cannot open snapshot file");
        fwrite(slice.data(), sizeof(float), slice.size(), f);
        fclose(f);
    }
};

struct FDTDKernel {
    // FDTDkernels.hpp:121-128
    int addRHS(FDTDGRIDS &, int itSample, FDTDMODELS &m, int &cb)
    {
        phc_check(phc_fd_add_rhs(m.dev, itSample, cb), "addRHS");
        return 0;
    }
    // FDTDkernels.hpp:75-84 (inner3D + applySponge in one launch)
    int computeOneStepSB(FDTDGRIDS &, FDTDMODELS &m, int &ca, int &cb)
    {
        phc_check(phc_fd_compute_one_step_sb(m.dev, ca, cb), "computeOneStepSB");
        return 0;
    }
    // FDTDkernels.hpp:86-104 (pml3D on the slabs, inner3D inside)
    int computeOneStepPML(FDTDGRIDS &, FDTDMODELS &m, int &ca, int &cb)
    {
        phc_check(phc_fd_compute_one_step_pml(m.dev, ca, cb), "computeOneStepPML");
        return 0;
    }
};

struct FDTDInit {
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
        m.coefx.assign(ncoefsX, 0.f);
        m.coefy.assign(ncoefsY, 0.f);
        m.coefz.assign(ncoefsZ, 0.f);
        phc_check(phc_fd_init_coef(g.lx, g.dx, m.coefx.data()), "init_coef x");
        phc_check(phc_fd_init_coef(g.ly, g.dy, m.coefy.data()), "init_coef y");
        phc_check(phc_fd_init_coef(g.lz, g.dz, m.coefz.data()), "init_coef z");
        m.coef0 = phc_fd_coef0(m.coefx.data(), g.lx, m.coefy.data(), g.ly, m.coefz.data(), g.lz);
        timeStep = phc_fd_compute_dt(vmax, m.coefx.data(), g.lx, m.coefy.data(), g.ly, m.coefz.data(),
                                     g.lz, cfl);
        nSamples = (int)(timeMax / timeStep);
        printf("init coefs done\n");
    }

    // FDTDinit.hpp:235-247
    void init_source(FDTDMODELS &m)
    {
        m.RHSTerm.assign(nSamples, 0.f);
        phc_check(phc_source_term(nSamples, timeStep, f0, sourceOrder, PHC_EXP_DOUBLE, m.RHSTerm.data()),
                  "source term");
    }

    // FDTDinit.hpp:249-317 (model) -- the device handle is created here
    void init_models(FDTDGRIDS &g, FDTDMODELS &m)
    {
        m.vp.assign((size_t)g.nx * g.ny * g.nz, 0.f);
        if (homogeneous) {
            const float v0 = ((vmin * vmin) * timeStep) * timeStep;
            for (float &v : m.vp) v = v0;
        } else if (layers > 0)
            phc_check(phc_fd_model_layered(g.nx, g.ny, g.nz, layers, vmin, vmax, layerSeed, timeStep,
                                           nullptr, m.vp.data()), "layered model");
        else
            phc_check(phc_fd_model_two_layer(g.nx, g.ny, g.nz, vmin, timeStep, m.vp.data()), "model");
    }

    // FDTDinit.hpp:319-396: the sponge is defined by its width and alpha and
    // evaluated on the device; this creates the device state.
    void defineSpongeBoundary(FDTDGRIDS &g, FDTDMODELS &m)
    {
        phc_fd_params p;
        std::memset(&p, 0, sizeof(p));
        p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
        p.lx = g.lx; p.ly = g.ly; p.lz = g.lz;
        p.xs = g.xs; p.ys = g.ys; p.zs = g.zs;
        p.coefx = m.coefx.data(); p.coefy = m.coefy.data(); p.coefz = m.coefz.data();
        p.coef0 = m.coef0;
        p.vp = m.vp.data();
        p.rhs = m.RHSTerm.data();
        p.nsamples = nSamples;
        p.sponge_width = usePML ? 0 : spongeWidth;  // computeOneStepPML applies no sponge
        p.sponge_alpha = spongeAlpha;
        p.init_value = 0.000001f;
        p.arith = arith;
        if (m.dev) phc_fd_destroy(m.dev);
        m.dev = nullptr;
        phc_check(phc_fd_create(&p, &m.dev), "phc_fd_create");
        if (usePML) {
            // FDTDutils.hpp:129-171 init_eta: per-axis quadratic profiles
            const int nd[3] = {g.ndampx, g.ndampy, g.ndampz};
            const int n[3] = {g.nx, g.ny, g.nz};
            const float d[3] = {g.dx, g.dy, g.dz};
            std::vector<float> prof[3];
            for (int a = 0; a < 3; ++a) {
                prof[a].assign(n[a] + 2, 0.f);
                if (nd[a] > 0)
                    phc_check(phc_fd_pml_profile(n[a], nd[a], phc_fd_pml_scale(timeStep, vmax, nd[a], d[a]),
                                                 prof[a].data()), "pml profile");
            }
            const float hd2[3] = {g.hdx_2, g.hdy_2, g.hdz_2};
            phc_check(phc_fd_set_pml(m.dev, nd, prof[0].data(), prof[1].data(), prof[2].data(), hd2), "set_pml");
        }
    }
};
