/*
 * sem_oracle.c -- CPU restatement of the reference SEM hot path and setup.
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 *
 * Compiled with -ffp-contract=off so the FP32 entry points round where the
 * reference's expressions round.  The element operator is the one of
 * SEMQkGLIntegrals.hpp:363-520 with the off-diagonal blocks B13/B23 written
 * with their correct indices (the reference's lines 488-492 and 503 index them
 * wrongly; on the axis-aligned boxes SEMmesh builds those blocks multiply
 * B entries below 1e-12 of the diagonal ones, so both give the same floats).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------- GLL */

static void legendre(int p, double x, double *P, double *dP)
{
    /* P_p(x) and P_p'(x) by the three-term recurrence */
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    if (p == 0) { *P = 1.0; *dP = 0.0; return; }
    for (int n = 2; n <= p; n++) {
        const double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
        const double d2 = d0 + (2.0 * n - 1.0) * p1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    }
    *P = p1;
    *dP = d1;
}

void orc_gll_points(int p, double *xi)
{
    /* endpoints and the roots of P_p' (SEMQkGLBasisFunctions.hpp:36-78 lists
     * their closed forms for p = 1..5); Newton from Chebyshev-Lobatto guesses,
     * with P_p'' from the Legendre ODE (1-x^2)P'' = 2xP' - p(p+1)P. */
    xi[0] = -1.0;
    xi[p] = 1.0;
    for (int i = 1; i < p; i++) {
        double x = -cos(M_PI * i / p);
        for (int it = 0; it < 100; it++) {
            double P, dP;
            legendre(p, x, &P, &dP);
            const double d2P = (2.0 * x * dP - p * (p + 1.0) * P) / (1.0 - x * x);
            const double dx = dP / d2P;
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        xi[i] = x;
    }
    /* symmetrise (the nodes are symmetric about 0) */
    for (int i = 0; i <= p / 2; i++) {
        const double a = 0.5 * (xi[p - i] - xi[i]);
        xi[i] = -a;
        xi[p - i] = a;
    }
    if (p % 2 == 0) xi[p / 2] = 0.0;
}

void orc_gll_weights(int p, double *w, int ref_rounding)
{
    double xi[16];
    orc_gll_points(p, xi);
    for (int i = 0; i <= p; i++) {
        double P, dP;
        legendre(p, xi[i], &P, &dP);
        w[i] = 2.0 / (p * (p + 1.0) * P * P);
        if (ref_rounding) w[i] = floor(w[i] * 1e8 + 0.5) / 1e8;
    }
}

void orc_gll_deriv(int p, const double *xi, double *dphi)
{
    /* l_j'(xi_i) at GLL nodes: (P_p(xi_i)/P_p(xi_j))/(xi_i - xi_j) off the
     * diagonal, -+p(p+1)/4 at the two end nodes, 0 elsewhere on the diagonal */
    double P[16], dP;
    for (int i = 0; i <= p; i++) legendre(p, xi[i], &P[i], &dP);
    for (int j = 0; j <= p; j++)
        for (int i = 0; i <= p; i++) {
            double v;
            if (i != j) v = (P[i] / P[j]) / (xi[i] - xi[j]);
            else if (i == 0) v = -p * (p + 1.0) / 4.0;
            else if (i == p) v = p * (p + 1.0) / 4.0;
            else v = 0.0;
            dphi[j * (p + 1) + i] = v;
        }
}

/* ------------------------------------------------------------------ mesh */

void orc_sem_mesh_init(orc_sem_mesh *m, int ex, int ey, int ez, float lx, float ly, float lz,
                       int order)
{
    /* SEMmesh.cpp:3-26 (3D: ZEROED2D = 1) */
    m->ex = ex; m->ey = ey; m->ez = ez;
    m->lx = lx; m->ly = ly; m->lz = lz;
    m->order = order;
    m->nx = ex * order + 1;
    m->ny = ey * order + 1;
    m->nz = ez * order + 1;
    m->hx = (float)((double)lx / (1. * ex));
    m->hy = (float)((double)ly / (1. * ey));
    m->hz = (float)((double)lz / (1. * ez));
}

void orc_sem_coord_1d(int order, int h, int n, float *coord)
{
    /* SEMmesh.cpp:81-139: element n of width h (an int in the reference) */
    double xid[16];
    orc_gll_points(order, xid);
    const float x0 = (float)(n * h);
    const float x1 = (float)((n + 1) * h);
    const float b = (float)((double)(x1 + x0) / 2.);
    const float a = b - x0;
    for (int j = 0; j <= order; j++) coord[j] = a * (float)xid[j] + b;
}

void orc_sem_nodes_list(const orc_sem_mesh *m, int *list)
{
    /* SEMmesh.cpp:205-231 */
    const int p = m->order, np = (p + 1) * (p + 1) * (p + 1);
    for (int j = 0; j < m->ey; j++)
        for (int k = 0; k < m->ez; k++)
            for (int i = 0; i < m->ex; i++) {
                const int n0 = i + k * m->ex + j * m->ex * m->ez;
                const int offset = i * p + k * p * m->nx + j * p * m->nx * m->nz;
                for (int mm = 0; mm <= p; mm++)
                    for (int n = 0; n <= p; n++)
                        for (int l = 0; l <= p; l++)
                            list[(int64_t)n0 * np + l + n * (p + 1) + mm * (p + 1) * (p + 1)] =
                                offset + l + n * m->nx + mm * m->nx * m->nz;
            }
}

int orc_sem_element_from_point(const orc_sem_mesh *m, float x, float y, float z)
{
    /* SEMmesh.cpp:266-285 */
    int eX = 0, eY = 0, eZ = 0;
    for (int i = 0; i < m->ex; i++)
        if (x >= i * m->hx && x < (i + 1) * m->hx) eX = i;
    for (int i = 0; i < m->ey; i++)
        if (y >= i * m->hy && y < (i + 1) * m->hy) eY = i;
    for (int i = 0; i < m->ez; i++)
        if (z >= i * m->hz && z < (i + 1) * m->hz) eZ = i;
    return eX + eZ * m->ex + eY * m->ex * m->ez;
}

int orc_sem_interior_nodes(const orc_sem_mesh *m, int *list)
{
    /* SEMmesh.cpp:314-344 (3D branch) */
    int c = 0;
    for (int k = 1; k < m->ny - 1; k++)
        for (int j = 1; j < m->nz - 1; j++)
            for (int i = 1; i < m->nx - 1; i++) list[c++] = i + j * m->nx + k * m->nx * m->nz;
    return c;
}

/* -------------------------------------------------- FP32 reference state */

struct orc_sem {
    orc_sem_mesh m;
    int p, np, nE, nN, nI;
    int *nodes;       /* [nE][np] */
    int *interior;    /* [nI] */
    float *model;     /* [nE] */
    float *B;         /* [nE][np][6] */
    float *mloc;      /* [nE][np] (before the division by c^2) */
    double w[16], D[256];
    float dt;
    float *M, *y;     /* [nN] */
    int threads;      /* > 1: colour-parallel element loop (orc_sem_set_threads) */
};

/* SEMQkGLIntegrals.hpp:206-359 computeB V2 (3D) for one element */
static void compute_b(const orc_sem *s, const float *X, const float *Y, const float *Z, float *B,
                      float *mloc)
{
    const int p = s->p, q = p + 1;
    const double *w = s->w, *D = s->D;
#define DPHI(a, b) D[(a) * q + (b)]
    for (int i3 = 0; i3 < q; i3++)
        for (int i2 = 0; i2 < q; i2++)
            for (int i1 = 0; i1 < q; i1++) {
                const int i = i1 + i2 * q + i3 * q * q;
                double j00 = 0, j01 = 0, j02 = 0, j10 = 0, j11 = 0, j12 = 0, j20 = 0, j21 = 0, j22 = 0;
                for (int a = 0; a < q; a++) {
                    const int j = a + i2 * q + i3 * q * q;
                    j00 += (double)X[j] * DPHI(a, i1);
                    j20 += (double)Y[j] * DPHI(a, i1);
                    j10 += (double)Z[j] * DPHI(a, i1);
                }
                for (int a = 0; a < q; a++) {
                    const int j = i1 + a * q + i3 * q * q;
                    j01 += (double)X[j] * DPHI(a, i2);
                    j21 += (double)Y[j] * DPHI(a, i2);
                    j11 += (double)Z[j] * DPHI(a, i2);
                }
                for (int a = 0; a < q; a++) {
                    const int j = i1 + i2 * q + a * q * q;
                    j02 += (double)X[j] * DPHI(a, i3);
                    j22 += (double)Y[j] * DPHI(a, i3);
                    j12 += (double)Z[j] * DPHI(a, i3);
                }
                const double detJ = fabs(j00 * (j11 * j22 - j21 * j12) - j01 * (j10 * j22 - j20 * j12) +
                                         j02 * (j10 * j21 - j20 * j11));
                const double i00 = j11 * j22 - j12 * j21, i01 = j02 * j21 - j01 * j22,
                             i02 = j01 * j12 - j02 * j11, i10 = j12 * j20 - j10 * j22,
                             i11 = j00 * j22 - j02 * j20, i12 = j02 * j10 - j00 * j12,
                             i20 = j10 * j21 - j11 * j20, i21 = j01 * j20 - j00 * j21,
                             i22 = j00 * j11 - j01 * j10;
                const double t00 = i00, t01 = i10, t02 = i20, t10 = i01, t11 = i11, t12 = i21,
                             t20 = i02, t21 = i12, t22 = i22;
                const double r = 1. / detJ;
                float *b = B + 6 * i;
                b[0] = (float)((i00 * t00 + i01 * t10 + i02 * t20) * r);
                b[1] = (float)((i10 * t01 + i11 * t11 + i12 * t21) * r);
                b[2] = (float)((i20 * t02 + i21 * t12 + i22 * t22) * r);
                b[3] = (float)((i00 * t01 + i01 * t11 + i02 * t21) * r);
                b[4] = (float)((i00 * t02 + i01 * t12 + i02 * t22) * r);
                b[5] = (float)((i10 * t02 + i11 * t12 + i12 * t22) * r);
                mloc[i] = (float)(w[i1] * w[i2] * w[i3] * detJ);
            }
#undef DPHI
}

/* SEMmesh.cpp:168-202: per-element coordinates (x, z, y local order) */
static void element_coords(const orc_sem_mesh *m, int e, float *X, float *Y, float *Z)
{
    const int p = m->order, q = p + 1;
    const int l = e % m->ex, mm = (e / m->ex) % m->ez, n = e / (m->ex * m->ez);
    float cx[16], cy[16], cz[16];
    orc_sem_coord_1d(p, (int)m->hy, n, cy);
    orc_sem_coord_1d(p, (int)m->hz, mm, cz);
    orc_sem_coord_1d(p, (int)m->hx, l, cx);
    for (int k = 0; k < q; k++)
        for (int j = 0; j < q; j++)
            for (int i = 0; i < q; i++) {
                X[i + q * j + k * q * q] = cx[i];
                Z[i + q * j + k * q * q] = cz[j];
                Y[i + q * j + k * q * q] = cy[k];
            }
}

orc_sem *orc_sem_create(const orc_sem_mesh *m, const float *model, float dt)
{
    orc_sem *s = (orc_sem *)calloc(1, sizeof(orc_sem));
    s->m = *m;
    s->p = m->order;
    const int q = s->p + 1;
    s->np = q * q * q;
    s->nE = m->ex * m->ey * m->ez;
    s->nN = m->nx * m->ny * m->nz;
    s->dt = dt;
    s->nodes = (int *)malloc(sizeof(int) * (size_t)s->nE * s->np);
    orc_sem_nodes_list(m, s->nodes);
    s->interior = (int *)malloc(sizeof(int) * (size_t)s->nN);
    s->nI = orc_sem_interior_nodes(m, s->interior);
    s->model = (float *)malloc(sizeof(float) * s->nE);
    memcpy(s->model, model, sizeof(float) * s->nE);
    double xi[16];
    orc_gll_points(s->p, xi);
    orc_gll_weights(s->p, s->w, 1);
    orc_gll_deriv(s->p, xi, s->D);
    s->B = (float *)malloc(sizeof(float) * 6 * (size_t)s->nE * s->np);
    s->mloc = (float *)malloc(sizeof(float) * (size_t)s->nE * s->np);
    s->M = (float *)calloc(s->nN, sizeof(float));
    s->y = (float *)calloc(s->nN, sizeof(float));
#pragma omp parallel
    {
        float X[216], Y[216], Z[216];
#pragma omp for schedule(static)
        for (int e = 0; e < s->nE; e++) {
            element_coords(m, e, X, Y, Z);
            compute_b(s, X, Y, Z, s->B + (size_t)6 * e * s->np, s->mloc + (size_t)e * s->np);
        }
    }
    return s;
}

void orc_sem_destroy(orc_sem *s)
{
    if (!s) return;
    free(s->nodes); free(s->interior); free(s->model); free(s->B); free(s->mloc);
    free(s->M); free(s->y);
    free(s);
}

int orc_sem_num_nodes(const orc_sem *s) { return s->nN; }

/* SEMQkGLIntegrals.hpp:363-520 gradPhiGradPhi (3D), one element: Y = R p */
static void grad_phi_grad_phi(const orc_sem *s, const float *B, const float *pl, float *Y)
{
    const int p = s->p, q = p + 1, q2 = q * q, np = s->np;
    const double *w = s->w, *D = s->D;
    float R[216];
#define DPHI(a, b) D[(a) * q + (b)]
#define BB(idx, c) ((double)B[6 * (idx) + (c)])
    for (int i3 = 0; i3 < q; i3++)
        for (int i2 = 0; i2 < q; i2++)
            for (int i1 = 0; i1 < q; i1++) {
                for (int j = 0; j < np; j++) R[j] = 0;
                for (int j1 = 0; j1 < q; j1++) {
                    const int j = j1 + i2 * q + i3 * q2;
                    for (int l = 0; l < q; l++) {
                        const int ll = l + i2 * q + i3 * q2;
                        R[j] = (float)((double)R[j] + w[l] * w[i2] * w[i3] *
                                                          (BB(ll, 0) * DPHI(i1, l) * DPHI(j1, l)));
                    }
                }
                for (int j2 = 0; j2 < q; j2++) {
                    const int j = i1 + j2 * q + i3 * q2;
                    for (int m = 0; m < q; m++) {
                        const int mm = i1 + m * q + i3 * q2;
                        R[j] = (float)((double)R[j] + w[i1] * w[m] * w[i3] *
                                                          (BB(mm, 1) * DPHI(i2, m) * DPHI(j2, m)));
                    }
                }
                for (int j3 = 0; j3 < q; j3++) {
                    const int j = i1 + i2 * q + j3 * q2;
                    for (int n = 0; n < q; n++) {
                        const int nn = i1 + i2 * q + n * q2;
                        R[j] = (float)((double)R[j] + w[i1] * w[i2] * w[n] *
                                                          (BB(nn, 2) * DPHI(i3, n) * DPHI(j3, n)));
                    }
                }
                for (int j2 = 0; j2 < q; j2++)
                    for (int j1 = 0; j1 < q; j1++) {
                        const int j = j1 + j2 * q + i3 * q2;
                        const int k = j1 + i2 * q + i3 * q2;
                        const int l = i1 + j2 * q + i3 * q2;
                        R[j] = (float)((double)R[j] +
                                       (w[j1] * w[i2] * w[i3] * (BB(k, 3) * DPHI(i1, j1) * DPHI(j2, i2)) +
                                        w[i1] * w[j2] * w[i3] * (BB(l, 3) * DPHI(j1, i1) * DPHI(i2, j2))));
                    }
                for (int j3 = 0; j3 < q; j3++)
                    for (int j1 = 0; j1 < q; j1++) {
                        const int j = j1 + i2 * q + j3 * q2;
                        const int k = j1 + i2 * q + i3 * q2;
                        const int l = i1 + i2 * q + j3 * q2;
                        R[j] = (float)((double)R[j] +
                                       (w[j1] * w[i2] * w[i3] * (BB(k, 4) * DPHI(i1, j1) * DPHI(j3, i3)) +
                                        w[i1] * w[i2] * w[j3] * (BB(l, 4) * DPHI(j1, i1) * DPHI(i3, j3))));
                    }
                for (int j3 = 0; j3 < q; j3++)
                    for (int j2 = 0; j2 < q; j2++) {
                        const int j = i1 + j2 * q + j3 * q2;
                        const int k = i1 + j2 * q + i3 * q2;
                        const int l = i1 + i2 * q + j3 * q2;
                        R[j] = (float)((double)R[j] +
                                       (w[i1] * w[j2] * w[i3] * (BB(k, 5) * DPHI(i2, j2) * DPHI(j3, i3)) +
                                        w[i1] * w[i2] * w[j3] * (BB(l, 5) * DPHI(j2, i2) * DPHI(i3, j3))));
                    }
                const int i = i1 + i2 * q + i3 * q2;
                float acc = 0;
                for (int j = 0; j < np; j++) acc = acc + R[j] * pl[j];
                Y[i] = acc;
            }
#undef DPHI
#undef BB
}

/* the element loop of SEMsolver.cpp:33-82 (K1 + K3), sequential order */
static void assemble(orc_sem *s, const float *p, float *y, float *M)
{
    const int np = s->np;
    for (int n = 0; n < s->nN; n++) { M[n] = 0; y[n] = 0; }
    float pl[216], Y[216];
    for (int e = 0; e < s->nE; e++) {
        const int *nl = s->nodes + (size_t)e * np;
        for (int i = 0; i < np; i++) pl[i] = p[nl[i]];
        grad_phi_grad_phi(s, s->B + (size_t)6 * e * np, pl, Y);
        const float c = s->model[e];
        const float *ml = s->mloc + (size_t)e * np;
        for (int i = 0; i < np; i++) {
            const int g = nl[i];
            const float mi = ml[i] / (c * c);
            M[g] += mi;
            y[g] += Y[i];
        }
    }
}

/* The element loop on several threads without the reference's race: the 8
 * colours of (x, z, y) element index parities run one after the other, and
 * the elements of one colour share no node, so their plain scatter adds
 * never meet (the 3D form of SEM_MESHCOLOR, SEMmesh.cpp:380-429).  Each
 * node's contributions arrive in colour order rather than element order, so
 * the FP32 sums can differ from the sequential loop's in the last bits:
 * a timing path (the CPU baseline), not the parity reference. */
static void assemble_colours(orc_sem *s, const float *p, float *y, float *M)
{
    const int np = s->np, ex = s->m.ex, ey = s->m.ey, ez = s->m.ez;
#pragma omp parallel for schedule(static) num_threads(s->threads)
    for (int n = 0; n < s->nN; n++) { M[n] = 0; y[n] = 0; }
#ifdef ORC_PLANT_SCATTER_RACE
    /* a planted mistake for the sanitizer check: all elements in one parallel
     * loop with plain adds, as the reference's OpenMP scatter does
     * (SEMmacros.hpp:50-56); ThreadSanitizer must report it */
    const int ncolours = 1;
#else
    const int ncolours = 8;
#endif
    for (int c = 0; c < ncolours; c++) {
        const int px = c & 1, pz = (c >> 1) & 1, py = (c >> 2) & 1;
        const int nx = (ex - px + 1) / 2, nz = (ez - pz + 1) / 2, ny = (ey - py + 1) / 2;
        const int ne = ncolours == 8 ? nx * nz * ny : s->nE;
#pragma omp parallel for schedule(static) num_threads(s->threads)
        for (int t = 0; t < ne; t++) {
            const int l = ncolours == 8 ? 2 * (t % nx) + px : t % ex;
            const int m = ncolours == 8 ? 2 * ((t / nx) % nz) + pz : (t / ex) % ez;
            const int n = ncolours == 8 ? 2 * (t / (nx * nz)) + py : t / (ex * ez);
            const int e = l + m * ex + n * ex * ez;
            float pl[216], Y[216];
            const int *nl = s->nodes + (size_t)e * np;
            for (int i = 0; i < np; i++) pl[i] = p[nl[i]];
            grad_phi_grad_phi(s, s->B + (size_t)6 * e * np, pl, Y);
            const float cc = s->model[e];
            const float *ml = s->mloc + (size_t)e * np;
            for (int i = 0; i < np; i++) {
                const int g = nl[i];
                M[g] += ml[i] / (cc * cc);
                y[g] += Y[i];
            }
        }
    }
}

void orc_sem_set_threads(orc_sem *s, int threads) { s->threads = threads; }

void orc_sem_apply_k(orc_sem *s, const float *p, float *y, float *M) { assemble(s, p, y, M); }

void orc_sem_step(orc_sem *s, float *p_i1, float *p_i2, int src_element, float rhs)
{
    /* SEMsolver.cpp:39-43: source into the first node of the source element */
    if (src_element >= 0) {
        const int node = s->nodes[(size_t)src_element * s->np];
        const float c = s->model[src_element];
        p_i2[node] += (((s->dt * s->dt) * c) * c) * rhs;
    }
    if (s->threads > 1)
        assemble_colours(s, p_i2, s->y, s->M);
    else
        assemble(s, p_i2, s->y, s->M);
    /* SEMsolver.cpp:85-89 */
#pragma omp parallel for schedule(static) num_threads(s->threads > 1 ? s->threads : 1)
    for (int n = 0; n < s->nI; n++) {
        const int I = s->interior[n];
        p_i1[I] = (2 * p_i2[I] - p_i1[I]) - ((s->dt * s->dt) * s->y[I]) / s->M[I];
    }
}

void orc_sem_get_my(const orc_sem *s, float *M, float *y)
{
    memcpy(M, s->M, sizeof(float) * s->nN);
    memcpy(y, s->y, sizeof(float) * s->nN);
}

/* ---------------------------------------------------------- FP64 oracle */

struct orc_sem64 {
    orc_sem_mesh m;
    int p, np, nE, nN, nI;
    int *nodes, *interior;
    double *model;
    double w[16], D[256];
    double jx, jy, jz;  /* half element sizes (exact box geometry) */
    double dt;
    double *M, *y;
};

orc_sem64 *orc_sem64_create(const orc_sem_mesh *m, const double *model, double dt)
{
    orc_sem64 *s = (orc_sem64 *)calloc(1, sizeof(orc_sem64));
    s->m = *m;
    s->p = m->order;
    const int q = s->p + 1;
    s->np = q * q * q;
    s->nE = m->ex * m->ey * m->ez;
    s->nN = m->nx * m->ny * m->nz;
    s->dt = dt;
    s->nodes = (int *)malloc(sizeof(int) * (size_t)s->nE * s->np);
    orc_sem_nodes_list(m, s->nodes);
    s->interior = (int *)malloc(sizeof(int) * (size_t)s->nN);
    s->nI = orc_sem_interior_nodes(m, s->interior);
    s->model = (double *)malloc(sizeof(double) * s->nE);
    memcpy(s->model, model, sizeof(double) * s->nE);
    double xi[16];
    orc_gll_points(s->p, xi);
    orc_gll_weights(s->p, s->w, 0);
    orc_gll_deriv(s->p, xi, s->D);
    s->jx = (double)m->lx / m->ex / 2.0;
    s->jy = (double)m->ly / m->ey / 2.0;
    s->jz = (double)m->lz / m->ez / 2.0;
    s->M = (double *)calloc(s->nN, sizeof(double));
    s->y = (double *)calloc(s->nN, sizeof(double));
    /* mass once: w w w detJ / c^2 */
    const double detJ = s->jx * s->jy * s->jz;
    for (int e = 0; e < s->nE; e++) {
        const int *nl = s->nodes + (size_t)e * s->np;
        const double c2 = model[e] * model[e];
        for (int i3 = 0; i3 < q; i3++)
            for (int i2 = 0; i2 < q; i2++)
                for (int i1 = 0; i1 < q; i1++)
                    s->M[nl[i1 + i2 * q + i3 * q * q]] += s->w[i1] * s->w[i2] * s->w[i3] * detJ / c2;
    }
    return s;
}

void orc_sem64_destroy(orc_sem64 *s)
{
    if (!s) return;
    free(s->nodes); free(s->interior); free(s->model); free(s->M); free(s->y);
    free(s);
}

/* y = K p, element by element, sum-factorised.  Local order (x, z, y); the
 * box Jacobian is diag(jx, jz, jy) in local axes, so B = diag(jz jy/jx,
 * jx jy/jz, jx jz/jy). */
void orc_sem64_apply_k(orc_sem64 *s, const double *p, double *y, double *M)
{
    const int q = s->p + 1, np = s->np;
    const double b1 = s->jz * s->jy / s->jx, b2 = s->jx * s->jy / s->jz, b3 = s->jx * s->jz / s->jy;
    const double *w = s->w, *D = s->D;
    for (int n = 0; n < s->nN; n++) y[n] = 0;
    double u[216], g1[216], g2[216], g3[216];
    for (int e = 0; e < s->nE; e++) {
        const int *nl = s->nodes + (size_t)e * np;
        for (int i = 0; i < np; i++) u[i] = p[nl[i]];
        for (int c = 0; c < q; c++)
            for (int b = 0; b < q; b++)
                for (int a = 0; a < q; a++) {
                    double s1 = 0, s2 = 0, s3 = 0;
                    for (int m = 0; m < q; m++) {
                        s1 += D[m * q + a] * u[m + b * q + c * q * q];
                        s2 += D[m * q + b] * u[a + m * q + c * q * q];
                        s3 += D[m * q + c] * u[a + b * q + m * q * q];
                    }
                    const double wq = w[a] * w[b] * w[c];
                    const int i = a + b * q + c * q * q;
                    g1[i] = wq * b1 * s1;
                    g2[i] = wq * b2 * s2;
                    g3[i] = wq * b3 * s3;
                }
        for (int c = 0; c < q; c++)
            for (int b = 0; b < q; b++)
                for (int a = 0; a < q; a++) {
                    double acc = 0;
                    for (int m = 0; m < q; m++) {
                        acc += D[a * q + m] * g1[m + b * q + c * q * q];
                        acc += D[b * q + m] * g2[a + m * q + c * q * q];
                        acc += D[c * q + m] * g3[a + b * q + m * q * q];
                    }
                    y[nl[a + b * q + c * q * q]] += acc;
                }
    }
    if (M) memcpy(M, s->M, sizeof(double) * s->nN);
}

void orc_sem64_step(orc_sem64 *s, double *p_i1, double *p_i2, int src_element, double rhs)
{
    if (src_element >= 0) {
        const int node = s->nodes[(size_t)src_element * s->np];
        const double c = s->model[src_element];
        p_i2[node] += s->dt * s->dt * c * c * rhs;
    }
    orc_sem64_apply_k(s, p_i2, s->y, NULL);
    for (int n = 0; n < s->nI; n++) {
        const int I = s->interior[n];
        p_i1[I] = 2 * p_i2[I] - p_i1[I] - s->dt * s->dt * s->y[I] / s->M[I];
    }
}

/* ---------------------------------------------- dense brute-force element */

static double lagrange_deriv(int p, const double *xi, int j, double x)
{
    /* d/dx prod_{m != j} (x - xi_m)/(xi_j - xi_m), by the product rule */
    double sum = 0.0;
    for (int k = 0; k <= p; k++) {
        if (k == j) continue;
        double term = 1.0 / (xi[j] - xi[k]);
        for (int m = 0; m <= p; m++) {
            if (m == j || m == k) continue;
            term *= (x - xi[m]) / (xi[j] - xi[m]);
        }
        sum += term;
    }
    return sum;
}

void orc_sem_dense_element_f64(int p, const double *X, const double *Y, const double *Z,
                               double *K, double *Mloc)
{
    /* K_ij = sum_q w_q |det J(q)| grad phi_i(q) . grad phi_j(q), with
     * grad = J^{-T} grad_xi, J computed from the node coordinates; phi are
     * tensor Lagrange polynomials on GLL nodes, quadrature on the same nodes */
    const int q = p + 1, np = q * q * q;
    double xi[16], w[16], Dd[16][16];
    orc_gll_points(p, xi);
    orc_gll_weights(p, w, 0);
    for (int j = 0; j < q; j++)
        for (int i = 0; i < q; i++) Dd[j][i] = lagrange_deriv(p, xi, j, xi[i]);
    for (int i = 0; i < np * np; i++) K[i] = 0;
    double *G = (double *)malloc(sizeof(double) * 3 * np);
    for (int c = 0; c < q; c++)
        for (int b = 0; b < q; b++)
            for (int a = 0; a < q; a++) {
                /* Jacobian d(x,y,z)/d(xi1,xi2,xi3) at quadrature node (a,b,c) */
                double J[3][3] = {{0}};
                for (int m = 0; m < q; m++) {
                    const int n1 = m + b * q + c * q * q, n2 = a + m * q + c * q * q,
                              n3 = a + b * q + m * q * q;
                    J[0][0] += X[n1] * Dd[m][a]; J[1][0] += Y[n1] * Dd[m][a]; J[2][0] += Z[n1] * Dd[m][a];
                    J[0][1] += X[n2] * Dd[m][b]; J[1][1] += Y[n2] * Dd[m][b]; J[2][1] += Z[n2] * Dd[m][b];
                    J[0][2] += X[n3] * Dd[m][c]; J[1][2] += Y[n3] * Dd[m][c]; J[2][2] += Z[n3] * Dd[m][c];
                }
                const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                   J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                /* inverse by cofactors: Jinv[r][s] */
                double Ji[3][3];
                Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
                Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
                Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
                Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
                Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
                Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
                Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
                Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
                Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
                /* physical gradient of every basis function at this node */
                for (int k3 = 0; k3 < q; k3++)
                    for (int k2 = 0; k2 < q; k2++)
                        for (int k1 = 0; k1 < q; k1++) {
                            const int n = k1 + k2 * q + k3 * q * q;
                            const double gx1 = Dd[k1][a] * (k2 == b) * (k3 == c);
                            const double gx2 = (k1 == a) * Dd[k2][b] * (k3 == c);
                            const double gx3 = (k1 == a) * (k2 == b) * Dd[k3][c];
                            for (int d = 0; d < 3; d++)
                                G[3 * n + d] = Ji[0][d] * gx1 + Ji[1][d] * gx2 + Ji[2][d] * gx3;
                        }
                const double wq = w[a] * w[b] * w[c] * fabs(det);
                for (int i = 0; i < np; i++)
                    for (int j = 0; j < np; j++)
                        K[i * np + j] += wq * (G[3 * i] * G[3 * j] + G[3 * i + 1] * G[3 * j + 1] +
                                               G[3 * i + 2] * G[3 * j + 2]);
                if (Mloc) Mloc[a + b * q + c * q * q] = wq;
            }
    free(G);
}
