/* sanitize_check.c -- drives every loop of the oracle at small sizes, for
 * sanitizer builds (oracle/sanitize.mk): AddressSanitizer + UndefinedBehavior
 * (out-of-bounds, use after free, leaks, overflow) and ThreadSanitizer with
 * the OpenMP runtime's race annotations (the FD loops, the PML step, the
 * SEM setup and the colour-parallel SEM element loop run on several
 * threads).  Test infrastructure; prints "SANITIZE OK" when every run ends.
 * The colour-parallel loop is the one that must not race: the reference's
 * OpenMP scatter does (SEMmacros.hpp:50-56, ATOMICADD without atomics). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int fd_case(int pml)
{
    orc_fd_grid g;
    memset(&g, 0, sizeof(g));
    g.nx = 22; g.ny = 19; g.nz = 26;
    g.lx = 4; g.ly = 3; g.lz = 2;
    g.xs = 11; g.ys = 9; g.zs = 13;
    g.x3 = 0; g.x4 = g.nx; g.y3 = 0; g.y4 = g.ny; g.z3 = 0; g.z4 = g.nz;
    float cx[7], cy[7], cz[7];
    orc_fd_init_coef(g.lx, 10.f, cx);
    orc_fd_init_coef(g.ly, 10.f, cy);
    orc_fd_init_coef(g.lz, 10.f, cz);
    const double coef0 = orc_fd_coef0(cx, g.lx, cy, g.ly, cz, g.lz);
    const float dt = orc_fd_compute_dt(4500.f, cx, g.lx, cy, g.ly, cz, g.lz, 0.8f);
    const int ns = 40;
    float *rhs = malloc(sizeof(float) * ns);
    orc_source_term(ns, dt, 25.f, 1, 1, rhs);
    const size_t n3 = (size_t)g.nx * g.ny * g.nz;
    const size_t ext = (size_t)(g.nx + 2 * g.lx) * (g.ny + 2 * g.ly) * (g.nz + 2 * g.lz);
    float *vp = malloc(sizeof(float) * n3), *sp = malloc(sizeof(float) * n3);
    for (size_t i = 0; i < n3; i++) vp[i] = 0.05f + 0.01f * (float)(i % 7);
    orc_fd_sponge(g.nx, g.ny, g.nz, 5, -0.00015f, sp);
    float *p0 = malloc(sizeof(float) * ext), *p1 = malloc(sizeof(float) * ext);
    for (size_t i = 0; i < ext; i++) p0[i] = p1[i] = 1e-6f;
    float tr[40];
    const int rec[6] = {0, 0, 0, 21, 18, 25};
    float rv[80];
    int cur = 1;
    if (pml) {
        const int nd[3] = {5, 4, 6};
        float *ex = calloc(g.nx + 2, sizeof(float)), *ey = calloc(g.ny + 2, sizeof(float)),
              *ez = calloc(g.nz + 2, sizeof(float));
        orc_fd_pml_profile(g.nx, nd[0], orc_fd_pml_scale(dt, 4500.f, nd[0], 10.f), ex);
        orc_fd_pml_profile(g.ny, nd[1], orc_fd_pml_scale(dt, 4500.f, nd[1], 10.f), ey);
        orc_fd_pml_profile(g.nz, nd[2], orc_fd_pml_scale(dt, 4500.f, nd[2], 10.f), ez);
        const float hd2[3] = {0.0025f, 0.0025f, 0.0025f};
        float *phi = calloc(n3, sizeof(float));
        orc_fd_run_pml(&g, nd, cx, cy, cz, coef0, vp, ex, ey, ez, hd2, phi, rhs, 0, ns, p0, p1, &cur, tr, 2, rec,
                       rv);
        free(ex); free(ey); free(ez); free(phi);
    } else {
        orc_fd_run(&g, cx, cy, cz, coef0, vp, sp, rhs, 0, ns, p0, p1, &cur, tr, 2, rec, rv);
    }
    const int ok = isfinite(tr[ns - 1]) && tr[ns - 1] != 0.f;
    free(rhs); free(vp); free(sp); free(p0); free(p1);
    return ok;
}

static int sem_case(int threads)
{
    orc_sem_mesh m;
    orc_sem_mesh_init(&m, 5, 4, 6, 100.f, 80.f, 120.f, 3);
    const int nE = m.ex * m.ey * m.ez;
    float *model = malloc(sizeof(float) * nE);
    for (int e = 0; e < nE; e++) model[e] = 1500.f + 37.f * (float)(e % 5);
    orc_sem *s = orc_sem_create(&m, model, 0.0004f);
    orc_sem_set_threads(s, threads);
    const int nN = orc_sem_num_nodes(s);
    float *a = calloc(nN, sizeof(float)), *b = calloc(nN, sizeof(float));
    const int src = orc_sem_element_from_point(&m, 51.f, 41.f, 61.f);
    float rhs[30];
    orc_source_term(30, 0.0004f, 25.f, 1, 0, rhs);
    for (int it = 0; it < 30; it++) {
        float *t;
        orc_sem_step(s, a, b, src, rhs[it]);
        t = a; a = b; b = t;
    }
    double e = 0;
    for (int n = 0; n < nN; n++) e += (double)b[n] * b[n];
    orc_sem_destroy(s);
    free(a); free(b); free(model);
    return isfinite(e) && e > 0;
}

static int sem2d_case(void)
{
    const int ex = 9, ez = 7;
    float model[63];
    for (int e = 0; e < ex * ez; e++) model[e] = 1500.f + 20.f * (float)(e % 3);
    orc_sem2d *s = orc_sem2d_create(ex, ez, 900.f, 700.f, 3, model, 0.0005f);
    int sz[7];
    orc_sem2d_sizes(s, sz);
    float *a = calloc(sz[2], sizeof(float)), *b = calloc(sz[2], sizeof(float));
    for (int it = 0; it < 40; it++) {
        float *t;
        orc_sem2d_step(s, a, b, 31, it < 20 ? 1.f : 0.f);
        t = a; a = b; b = t;
    }
    const int ok = isfinite(b[sz[2] / 2]);
    orc_sem2d_destroy(s);
    free(a); free(b);
    return ok;
}

int main(void)
{
    int ok = fd_case(0) && fd_case(1) && sem_case(1) && sem_case(4) && sem2d_case();
    printf(ok ? "SANITIZE OK\n" : "SANITIZE FAILED\n");
    return ok ? 0 : 1;
}
