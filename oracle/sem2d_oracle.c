/*
 * sem2d_oracle.c -- CPU restatement of the reference's SEM2D step (the
 * reference built with SEM2D: quadrilateral elements on the x-z plane, a
 * first-order paraxial absorbing boundary on all four edges).
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 *
 *   mesh sizes, ey = 0                  src/sem/mesh/SEMmesh.cpp:3-26, 28-55
 *   node coordinates                    SEMmesh.cpp:81-166 (see below)
 *   element node list (2D)              SEMmesh.cpp:205-231
 *   interior node list (2D branch)      SEMmesh.cpp:314-329
 *   boundary faces / nodes / face map   SEMmesh.cpp:433-622
 *   computeB (2D branch)                src/sem/finiteElements/SEMQkGLIntegrals.hpp:55-102
 *   gradPhiGradPhi (2D branch)          SEMQkGLIntegrals.hpp:372-425
 *   computeDs                           SEMQkGLIntegrals.hpp:568-599
 *   computeOneStep incl. the SEM2D part src/sem/solver/SEMsolver.cpp:21-125
 *
 * Coordinates.  The checked-in revision fills neither coordinate array for a
 * 2D mesh: the per-element arrays computeB V2 reads are built by a loop over
 * ey elements (SEMmesh.cpp:175-199, empty when ey = 0), and the global array
 * computeB V1 and computeDs read is filled from one element only
 * (getCoordInOneDirection's element loop is commented out, SEMmesh.cpp:128-137).
 * Built as it stands the 2D step divides by a zero Jacobian.  This restatement
 * gives every node the coordinate that loop, restored, assigns
 * (element n of width h, h the int element size, later elements overwriting
 * shared end nodes) and feeds it to the 2D formulas of computeB V1, which is
 * what the published 2D runs computed.
 *
 * Compiled with -ffp-contract=off; FP32/FP64 mixing follows the reference's
 * expressions.  Element contributions are added in element order (the
 * sequential build's order).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_sem2d {
    int ex, ez, p, q, nx, nz, nE, nN, nI, nB, nF;
    int *nodes;      /* [nE][q*q] */
    int *interior;   /* [nI] */
    int *bnodes;     /* [nB] boundary node ids, SEMmesh.cpp:499-531 order */
    int *face;       /* [nF][2 + q]: element, face id, node ids (faceInfos) */
    int *fmap;       /* [nF][q]: face node -> boundary-node index */
    float *model;    /* [nE] */
    float *cx, *cz;  /* node coordinates per axis */
    float *B;        /* [nE][q*q][4] */
    float *mloc;     /* [nE][q*q] */
    double w[16], D[256];
    float dt;
    int paraxial;    /* 0: boundary term left out (natural, reflecting edge) */
    float *M, *y, *Sh;
};

/* SEMmesh.cpp:442-497 getBoundaryFacesInfos: bottom (face id 1), right (2),
 * top (3), left (0) */
static void face_infos(orc_sem2d *s)
{
    const int p = s->p, q = s->q, ex = s->ex, ez = s->ez, nx = s->nx, st = 2 + q;
    for (int i = 0; i < ex; i++) {
        int *f = s->face + (size_t)i * st;
        f[0] = i;
        f[1] = 1;
        for (int j = 0; j <= p; j++) f[2 + j] = i * p + j;
    }
    for (int j = 0; j < ez; j++) {
        int *f = s->face + (size_t)(ex + j) * st;
        const int off = (ex - 1) * p + j * p * nx;
        f[0] = ex - 1 + j * ex;
        f[1] = 2;
        for (int k = 0; k <= p; k++) f[2 + k] = off + p + k * nx;
    }
    for (int i = 0; i < ex; i++) {
        int *f = s->face + (size_t)(ez + ex + i) * st;
        const int off = i * p + (ez - 1) * p * nx;
        f[0] = i + (ez - 1) * ex;
        f[1] = 3;
        for (int l = 0; l <= p; l++) f[2 + l] = off + l + p * nx;
    }
    for (int j = 0; j < ez; j++) {
        int *f = s->face + (size_t)(2 * ex + ez + j) * st;
        f[0] = j * ex;
        f[1] = 0;
        for (int k = 0; k <= p; k++) f[2 + k] = j * p * nx + k * nx;
    }
}

/* SEMmesh.cpp:499-531 getListOfBoundaryNodes */
static void boundary_nodes(orc_sem2d *s)
{
    const int nx = s->nx, nz = s->nz;
    int k = 0;
    for (int i = 0; i < nx; i++) s->bnodes[k++] = i;
    for (int j = 1; j < nz; j++) s->bnodes[k++] = nx - 1 + j * nx;
    for (int i = 0; i < nx - 1; i++) s->bnodes[k++] = i + (nz - 1) * nx;
    for (int j = 1; j < nz - 1; j++) s->bnodes[k++] = j * nx;
}

/* SEMmesh.cpp:533-622 getLocalFaceNodeToGlobalFaceNode: the index, in the
 * boundary-node list, of each face node, built face after face from the
 * previous face's last entry */
static void face_map(orc_sem2d *s)
{
    const int p = s->p, q = s->q, ex = s->ex, ez = s->ez;
#define FM(f, i) s->fmap[(size_t)(f) * q + (i)]
    for (int i = 0; i < ex; i++)
        for (int j = 0; j <= p; j++) FM(i, j) = i * p + j;
    for (int j = 0; j < ez; j++) {
        const int f = ex + j, off = FM(f - 1, p);
        for (int k = 0; k <= p; k++) FM(f, k) = off + k;
    }
    for (int i = 0; i < ex; i++) {
        const int f = ez + ex + i;
        if (i < ex - 1) {
            const int off = (i == 0) ? FM(f - 1, p) + 1 : FM(f - 1, p);
            for (int l = 0; l <= p; l++) FM(f, l) = off + l;
        } else {
            const int off = FM(f - 1, p);
            for (int l = 0; l < p; l++) FM(f, l) = off + l;
            FM(f, p) = FM(ex + ez - 1, p);
        }
    }
    for (int j = 0; j < ez; j++) {
        const int f = 2 * ex + ez + j;
        if (j == 0) {
            FM(f, 0) = 0;
            for (int k = 1; k <= p; k++) FM(f, k) = FM(f - 1, p - 1) + k;
        } else if (j < ez - 1) {
            for (int k = 0; k <= p; k++) FM(f, k) = FM(f - 1, p) + k;
        } else {
            for (int k = 0; k < p; k++) FM(f, k) = FM(f - 1, p) + k;
            FM(f, p) = FM(ex + ez, 0);
        }
    }
#undef FM
}

/* SEMQkGLIntegrals.hpp:55-102 computeB, 2D branch, for one element */
static void compute_b2d(const orc_sem2d *s, const int *nl, float *B, float *mloc)
{
    const int q = s->q;
    const double *w = s->w, *D = s->D;
#define DPHI(a, b) D[(a) * q + (b)]
#define XC(node) ((double)s->cx[(node) % s->nx])
#define ZC(node) ((double)s->cz[(node) / s->nx])
    for (int i2 = 0; i2 < q; i2++)
        for (int i1 = 0; i1 < q; i1++) {
            double j0 = 0, j1 = 0, j2 = 0, j3 = 0;
            const int i = i1 + i2 * q;
            for (int a = 0; a < q; a++) {
                const int g = nl[a + i2 * q];
                j0 += XC(g) * DPHI(a, i1);
                j2 += ZC(g) * DPHI(a, i1);
            }
            for (int a = 0; a < q; a++) {
                const int g = nl[i1 + a * q];
                j1 += XC(g) * DPHI(a, i2);
                j3 += ZC(g) * DPHI(a, i2);
            }
            const double det = fabs(j0 * j3 - j2 * j1);
            const double a0 = j3, a1 = -j1, a2 = -j2, a3 = j0;  /* adjugate */
            const double t0 = j3, t1 = -j2, t2 = -j1, t3 = j0;  /* its transpose */
            const double r = 1. / det;
            B[4 * i + 0] = (float)((a0 * t0 + a1 * t2) * r);
            B[4 * i + 1] = (float)((a0 * t1 + a1 * t3) * r);
            B[4 * i + 2] = (float)((a2 * t0 + a3 * t2) * r);
            B[4 * i + 3] = (float)((a2 * t1 + a3 * t3) * r);
            mloc[i] = (float)(w[i1] * w[i2] * det);
        }
#undef XC
#undef ZC
}

/* SEMQkGLIntegrals.hpp:372-425 gradPhiGradPhi, 2D branch: Y = R p per output
 * node (the cross blocks index B at (i1, j2) and (i2, j1) as written there) */
static void grad_phi_grad_phi_2d(const orc_sem2d *s, const float *B, const float *pl, float *Y)
{
    const int q = s->q, np = q * q;
    const double *w = s->w, *D = s->D;
    float R[36];
#define BB(idx, c) ((double)B[4 * (idx) + (c)])
    for (int i2 = 0; i2 < q; i2++)
        for (int i1 = 0; i1 < q; i1++) {
            for (int j = 0; j < np; j++) R[j] = 0;
            for (int j1 = 0; j1 < q; j1++) {
                const int j = j1 + i2 * q;
                for (int m = 0; m < q; m++)
                    R[j] = (float)((double)R[j] +
                                   w[m] * w[i2] * (BB(m + i2 * q, 0) * DPHI(i1, m) * DPHI(j1, m)));
            }
            for (int j1 = 0; j1 < q; j1++)
                for (int j2 = 0; j2 < q; j2++) {
                    const int j = j1 + j2 * q;
                    R[j] = (float)((double)R[j] +
                                   w[i1] * w[j2] * (BB(i1 + j2 * q, 1) * DPHI(i2, j2) * DPHI(j1, i1)));
                }
            for (int j1 = 0; j1 < q; j1++)
                for (int j2 = 0; j2 < q; j2++) {
                    const int j = j1 + j2 * q;
                    R[j] = (float)((double)R[j] +
                                   w[i2] * w[j1] * (BB(i2 + j1 * q, 2) * DPHI(i1, j1) * DPHI(j2, i2)));
                }
            for (int j2 = 0; j2 < q; j2++) {
                const int j = i1 + j2 * q;
                for (int n = 0; n < q; n++)
                    R[j] = (float)((double)R[j] +
                                   w[i1] * w[n] * (BB(i1 + n * q, 3) * DPHI(i2, n) * DPHI(j2, n)));
            }
            float acc = 0;
            for (int j = 0; j < np; j++) acc = acc + R[j] * pl[j];
            Y[i1 + i2 * q] = acc;
        }
#undef BB
#undef DPHI
}

orc_sem2d *orc_sem2d_create(int ex, int ez, float lx, float lz, int order, const float *model, float dt)
{
    if (ex < 2 || ez < 2 || order < 1 || order > 5) return NULL;
    orc_sem2d *s = (orc_sem2d *)calloc(1, sizeof(orc_sem2d));
    const int p = order, q = p + 1;
    s->ex = ex; s->ez = ez; s->p = p; s->q = q;
    s->nx = ex * p + 1;
    s->nz = ez * p + 1;
    s->nE = ex * ez;
    s->nN = s->nx * s->nz;
    s->nI = (s->nx - 2) * (s->nz - 2);
    s->nB = 2 * (s->nx + s->nz) - 4;  /* SEMmesh.cpp:437-439 */
    s->nF = 2 * (ex + ez);           /* SEMmesh.cpp:433-435 */
    s->dt = dt;
    s->paraxial = 1;
    s->nodes = (int *)malloc(sizeof(int) * (size_t)s->nE * q * q);
    /* SEMmesh.cpp:205-231 with ey = 0: one layer, node = offset + l + n nx */
    for (int k = 0; k < ez; k++)
        for (int i = 0; i < ex; i++) {
            const int e = i + k * ex, off = i * p + k * p * s->nx;
            for (int n = 0; n < q; n++)
                for (int l = 0; l < q; l++) s->nodes[(size_t)e * q * q + l + n * q] = off + l + n * s->nx;
        }
    s->interior = (int *)malloc(sizeof(int) * (size_t)s->nI);
    {
        int m = 0;  /* SEMmesh.cpp:318-328 */
        for (int j = 1; j < s->nz - 1; j++)
            for (int i = 1; i < s->nx - 1; i++) s->interior[m++] = i + j * s->nx;
    }
    s->bnodes = (int *)malloc(sizeof(int) * (size_t)s->nB);
    s->face = (int *)malloc(sizeof(int) * (size_t)s->nF * (2 + q));
    s->fmap = (int *)malloc(sizeof(int) * (size_t)s->nF * q);
    boundary_nodes(s);
    face_infos(s);
    face_map(s);
    s->model = (float *)malloc(sizeof(float) * s->nE);
    memcpy(s->model, model, sizeof(float) * s->nE);
    /* element sizes: float quotients (SEMmesh.cpp:20-22) handed to the
     * coordinate routine as ints (SEMmesh.hpp:61) */
    const int hx = (int)(float)((double)lx / (1. * ex)), hz = (int)(float)((double)lz / (1. * ez));
    s->cx = (float *)malloc(sizeof(float) * s->nx);
    s->cz = (float *)malloc(sizeof(float) * s->nz);
    float c[16];
    for (int e = 0; e < ex; e++) {
        orc_sem_coord_1d(p, hx, e, c);
        for (int j = 0; j < q; j++) s->cx[e * p + j] = c[j];
    }
    for (int e = 0; e < ez; e++) {
        orc_sem_coord_1d(p, hz, e, c);
        for (int j = 0; j < q; j++) s->cz[e * p + j] = c[j];
    }
    double xi[16];
    orc_gll_points(p, xi);
    orc_gll_weights(p, s->w, 1);
    orc_gll_deriv(p, xi, s->D);
    s->B = (float *)malloc(sizeof(float) * 4 * (size_t)s->nE * q * q);
    s->mloc = (float *)malloc(sizeof(float) * (size_t)s->nE * q * q);
    for (int e = 0; e < s->nE; e++)
        compute_b2d(s, s->nodes + (size_t)e * q * q, s->B + (size_t)4 * e * q * q, s->mloc + (size_t)e * q * q);
    s->M = (float *)calloc(s->nN, sizeof(float));
    s->y = (float *)calloc(s->nN, sizeof(float));
    s->Sh = (float *)calloc(s->nB, sizeof(float));
    return s;
}

void orc_sem2d_destroy(orc_sem2d *s)
{
    if (!s) return;
    free(s->nodes); free(s->interior); free(s->bnodes); free(s->face); free(s->fmap);
    free(s->model); free(s->cx); free(s->cz); free(s->B); free(s->mloc);
    free(s->M); free(s->y); free(s->Sh);
    free(s);
}

void orc_sem2d_sizes(const orc_sem2d *s, int *sizes)
{
    sizes[0] = s->nx; sizes[1] = s->nz; sizes[2] = s->nN; sizes[3] = s->nE;
    sizes[4] = s->nI; sizes[5] = s->nB; sizes[6] = s->nF;
}

void orc_sem2d_lists(const orc_sem2d *s, int *nodes, int *interior, int *bnodes, int *face, int *fmap)
{
    const int q = s->q;
    if (nodes) memcpy(nodes, s->nodes, sizeof(int) * (size_t)s->nE * q * q);
    if (interior) memcpy(interior, s->interior, sizeof(int) * (size_t)s->nI);
    if (bnodes) memcpy(bnodes, s->bnodes, sizeof(int) * (size_t)s->nB);
    if (face) memcpy(face, s->face, sizeof(int) * (size_t)s->nF * (2 + q));
    if (fmap) memcpy(fmap, s->fmap, sizeof(int) * (size_t)s->nF * q);
}

void orc_sem2d_coords(const orc_sem2d *s, float *cx, float *cz)
{
    memcpy(cx, s->cx, sizeof(float) * s->nx);
    memcpy(cz, s->cz, sizeof(float) * s->nz);
}

/* SEMsolver.cpp:33-36 + 46-82 (element loop), 2D */
static void assemble2d(orc_sem2d *s, const float *p, float *y, float *M)
{
    const int np = s->q * s->q;
    for (int n = 0; n < s->nN; n++) { M[n] = 0; y[n] = 0; }
    float pl[36], Y[36];
    for (int e = 0; e < s->nE; e++) {
        const int *nl = s->nodes + (size_t)e * np;
        for (int i = 0; i < np; i++) pl[i] = p[nl[i]];
        grad_phi_grad_phi_2d(s, s->B + (size_t)4 * e * np, pl, Y);
        const float c = s->model[e];
        const float *ml = s->mloc + (size_t)e * np;
        for (int i = 0; i < np; i++) {
            const int g = nl[i];
            M[g] += ml[i] / (c * c);
            y[g] += Y[i];
        }
    }
}

/* SEMQkGLIntegrals.hpp:568-599 computeDs + SEMsolver.cpp:95-110: boundary
 * term Sh per boundary node, face after face */
static void boundary_term(orc_sem2d *s)
{
    const int p = s->p, q = s->q, st = 2 + q;
    const double *D = s->D;
    for (int i = 0; i < s->nB; i++) s->Sh[i] = 0;
    for (int f = 0; f < s->nF; f++) {
        const int *fi = s->face + (size_t)f * st;
        float ds[6];
        for (int j = 0; j <= p; j++) {
            float js0 = 0, js1 = 0;
            for (int i = 0; i <= p; i++) {
                const int g = fi[2 + i];
                const float xi = s->cx[g % s->nx], zi = s->cz[g / s->nx];
                /* both branches of the reference (faces 0/2 and 1/3) are the same sum */
                js0 = (float)((double)js0 + D[i * q + j] * xi);
                js1 = (float)((double)js1 + D[i * q + j] * zi);
            }
            ds[j] = sqrtf(js0 * js0 + js1 * js1);
        }
        if (!s->paraxial) continue;
        for (int i = 0; i <= p; i++) {
            const float sh = (float)(s->w[i] * ds[i] / s->model[fi[0]]);
            s->Sh[s->fmap[(size_t)f * q + i]] += sh;
        }
    }
}

void orc_sem2d_apply_k(orc_sem2d *s, const float *p, float *y, float *M) { assemble2d(s, p, y, M); }

void orc_sem2d_step(orc_sem2d *s, float *p_i1, float *p_i2, int src_element, float rhs)
{
    const int np = s->q * s->q;
    /* SEMsolver.cpp:39-43 */
    if (src_element >= 0) {
        const int node = s->nodes[(size_t)src_element * np];
        const float c = s->model[src_element];
        p_i2[node] += (((s->dt * s->dt) * c) * c) * rhs;
    }
    assemble2d(s, p_i2, s->y, s->M);
    /* SEMsolver.cpp:85-89 */
    for (int n = 0; n < s->nI; n++) {
        const int I = s->interior[n];
        p_i1[I] = (2 * p_i2[I] - p_i1[I]) - ((s->dt * s->dt) * s->y[I]) / s->M[I];
    }
    /* SEMsolver.cpp:91-123 */
    boundary_term(s);
    for (int i = 0; i < s->nB; i++) {
        const int I = s->bnodes[i];
        const float M = s->M[I];
        const float invMpSh = (float)(1 / ((double)M + (double)(s->dt * s->Sh[i]) * 0.5));
        const float MmSh = (float)((double)M - (double)(s->dt * s->Sh[i]) * 0.5);
        p_i1[I] = invMpSh * (((2 * M) * p_i2[I] - MmSh * p_i1[I]) - ((s->dt * s->dt) * s->y[I]));
    }
}

void orc_sem2d_get_my(const orc_sem2d *s, float *M, float *y, float *Sh)
{
    if (M) memcpy(M, s->M, sizeof(float) * s->nN);
    if (y) memcpy(y, s->y, sizeof(float) * s->nN);
    if (Sh) memcpy(Sh, s->Sh, sizeof(float) * s->nB);
}

void orc_sem2d_boundary_term(orc_sem2d *s, float *Sh)
{
    boundary_term(s);
    memcpy(Sh, s->Sh, sizeof(float) * s->nB);
}

void orc_sem2d_set_paraxial(orc_sem2d *s, int on) { s->paraxial = on ? 1 : 0; }

/* SEMmesh.cpp:380-429 getNumberOfElementsByColor / sortElementsByColor (the
 * reference's 4 colours, 2D): red (even i, even j), green (odd i, even j),
 * blue (even i, odd j), yellow (odd i, odd j), element i + j ex, j outer */
int orc_sem2d_colors(int ex, int ez, int *counts, int *lists)
{
    const int mpc = (ex / 2 + ex % 2) * (ez / 2 + ez % 2);
    const int i0[4] = {0, 1, 0, 1}, j0[4] = {0, 0, 1, 1};
    for (int c = 0; c < 4; c++) {
        int k = 0;
        for (int j = j0[c]; j < ez; j = j + 2)
            for (int i = i0[c]; i < ex; i = i + 2) {
                lists[c * mpc + k] = i + j * ex;
                k = k + 1;
            }
        counts[c] = k;
    }
    return mpc;
}

/* ------------------------------------------------------------- FP64 2D */
/* The same 2D step in FP64 on exact box geometry (J = diag(hx/2, hz/2), the
 * int element sizes), the reference's 8-decimal GLL weights, double fields:
 * the yardstick of the FP32 paths' rounding. */
struct orc_sem2d64 {
    int ex, ez, p, q, nx, nz, nE, nN;
    double hx, hz, dt;
    double w[16], D[256];
    double *model, *M, *Sh, *y;
    int src_node;
    double src_coef;
};

orc_sem2d64 *orc_sem2d64_create(int ex, int ez, float lx, float lz, int order, const double *model, double dt,
                                int src_element)
{
    if (ex < 2 || ez < 2 || order < 1 || order > 5) return NULL;
    orc_sem2d64 *s = (orc_sem2d64 *)calloc(1, sizeof(orc_sem2d64));
    const int p = order, q = p + 1;
    s->ex = ex; s->ez = ez; s->p = p; s->q = q;
    s->nx = ex * p + 1; s->nz = ez * p + 1;
    s->nE = ex * ez; s->nN = s->nx * s->nz;
    s->hx = (int)(float)((double)lx / (1. * ex));
    s->hz = (int)(float)((double)lz / (1. * ez));
    s->dt = dt;
    double xi[16];
    orc_gll_points(p, xi);
    orc_gll_weights(p, s->w, 1);
    orc_gll_deriv(p, xi, s->D);
    s->model = (double *)malloc(sizeof(double) * s->nE);
    memcpy(s->model, model, sizeof(double) * s->nE);
    s->M = (double *)calloc(s->nN, sizeof(double));
    s->Sh = (double *)calloc(s->nN, sizeof(double));
    s->y = (double *)calloc(s->nN, sizeof(double));
    for (int k = 0; k < ez; k++)
        for (int i = 0; i < ex; i++) {
            const double c = model[i + k * ex];
            for (int b = 0; b < q; b++)
                for (int a = 0; a < q; a++)
                    s->M[(i * p + a) + (k * p + b) * s->nx] += s->w[a] * s->w[b] * s->hx * s->hz / 4 / (c * c);
            for (int a = 0; a < q; a++) {
                if (k == 0) s->Sh[i * p + a] += s->w[a] * s->hx / 2 / c;
                if (k == ez - 1) s->Sh[i * p + a + (s->nz - 1) * s->nx] += s->w[a] * s->hx / 2 / c;
                if (i == 0) s->Sh[(k * p + a) * s->nx] += s->w[a] * s->hz / 2 / c;
                if (i == ex - 1) s->Sh[(k * p + a) * s->nx + s->nx - 1] += s->w[a] * s->hz / 2 / c;
            }
        }
    const int e = src_element, i0 = e % ex, k0 = e / ex;
    s->src_node = i0 * p + k0 * p * s->nx;
    s->src_coef = dt * dt * model[e] * model[e];
    return s;
}

void orc_sem2d64_destroy(orc_sem2d64 *s)
{
    if (!s) return;
    free(s->model); free(s->M); free(s->Sh); free(s->y);
    free(s);
}

void orc_sem2d64_step(orc_sem2d64 *s, double *p_i1, double *p_i2, double rhs)
{
    const int p = s->p, q = s->q, nx = s->nx;
    const double *w = s->w, *D = s->D;
    const double bx = s->hz / s->hx, bz = s->hx / s->hz;  /* |J| / J_xx^2, |J| / J_zz^2 */
    p_i2[s->src_node] += s->src_coef * rhs;
    for (int n = 0; n < s->nN; n++) s->y[n] = 0;
    for (int k = 0; k < s->ez; k++)
        for (int i = 0; i < s->ex; i++)
            for (int i2 = 0; i2 < q; i2++)
                for (int i1 = 0; i1 < q; i1++) {
                    double acc = 0;
                    for (int j1 = 0; j1 < q; j1++) {
                        double r = 0;
                        for (int m = 0; m < q; m++) r += w[m] * D[i1 * q + m] * D[j1 * q + m];
                        acc += w[i2] * bx * r * p_i2[(i * p + j1) + (k * p + i2) * nx];
                    }
                    for (int j2 = 0; j2 < q; j2++) {
                        double r = 0;
                        for (int m = 0; m < q; m++) r += w[m] * D[i2 * q + m] * D[j2 * q + m];
                        acc += w[i1] * bz * r * p_i2[(i * p + i1) + (k * p + j2) * nx];
                    }
                    s->y[(i * p + i1) + (k * p + i2) * nx] += acc;
                }
    for (int n = 0; n < s->nN; n++) {
        const double a = s->dt * s->Sh[n] * 0.5;
        p_i1[n] = (2 * s->M[n] * p_i2[n] - (s->M[n] - a) * p_i1[n] - s->dt * s->dt * s->y[n]) / (s->M[n] + a);
    }
}
