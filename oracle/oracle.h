/*
 * oracle.h -- CPU restatement of the reference's time-stepping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libphc.so, the C++ host
 * drivers, the kernels) links, loads or calls this code.  It is the checker
 * that the HIP path is compared against: pytest, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py are its only users.
 *
 * Every function names the reference file:line it restates (paths relative
 * to the numpex/proxy-geos-hc tree).  The arithmetic follows the reference
 * operation by operation -- same types, same evaluation order, no FMA
 * contraction (this file is compiled with -ffp-contract=off) -- so the FP32
 * "reference semantics" entry points reproduce the reference's own rounding.
 * The *_f64 entry points are an all-FP64 restatement used as the precision
 * yardstick.
 *
 * Parity pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_fd_init_coef / orc_fd_coef0 / orc_fd_compute_dt : closed forms and the
 *       published step-0 value 0.041887 (test_20240512 FD logs).
 *   orc_ricker / orc_source_term : FP64 closed form; the published SEM
 *       "sourceTerm = 0.0102097" lines.
 *   orc_fd_sponge : 20-entry table exp(alpha (L-i)^2) and the face-slab rule.
 *   orc_fd_run    : the 25 published FD source-point prints (150^3, 1242
 *       steps) of the reference's CPU builds; Laplacian of x^2+y^2+z^2 = 6;
 *       brute-force FP64 triple loop.
 *   orc_gll_*     : roots of P'_p, w = 2/(p(p+1)P_p^2), the reference's
 *       8-digit weight table, row sums of D = 0, exact polynomial derivatives.
 *   orc_sem_*     : K 1 = 0, K symmetric positive semi-definite, quadratic
 *       exactness, brute-force dense FP64 K on a 2x2x2 mesh, Sum M = volume/c^2.
 */
#ifndef PHC_ORACLE_H
#define PHC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- source */
/* utils.hpp:11-46 evaluateRicker.  use_double_exp selects ::exp(double) (the FD
 * translation unit) or std::exp(float) (the SEM translation unit). */
float orc_ricker(float time_n, float f0, int order, int use_double_exp);
/* utils.hpp:48-57 computeSourceTerm */
void orc_source_term(int nsamples, float dt, float f0, int order,
                     int use_double_exp, float *out);
/* FP64 closed form of the same pulse (precision yardstick). */
double orc_ricker_f64(double t, double f0, int order);

/* -------------------------------------------------------------------- FD */
/* FDTDutils.hpp:13-58 init_coef: coef[0..L] */
void orc_fd_init_coef(int L, float dx, float *coef);
/* FDTDinit.hpp:212-227: coef0 = -2 (sum_x + sum_y + sum_z), sums in float */
double orc_fd_coef0(const float *cx, int Lx, const float *cy, int Ly,
                    const float *cz, int Lz);
/* FDTDutils.hpp:59-76 compute_dt_sch (cfl is 0.8 in the reference) */
float orc_fd_compute_dt(float vmax, const float *cx, int Lx, const float *cy,
                        int Ly, const float *cz, int Lz, float cfl);
/* FDTDinit.hpp:319-396 defineSpongeBoundary: full nx*ny*nz array, IDX3 order */
void orc_fd_sponge(int nx, int ny, int nz, int L, float alpha, float *sponge);
/* sponge value table exp(alpha*(L-i)^2), i = 0..L-1 (the per-cell factor) */
void orc_fd_sponge_table(int L, float alpha, float *tab);

typedef struct {
    int nx, ny, nz;       /* interior points (FDTDdata.hpp:8-9) */
    int lx, ly, lz;       /* stencil half widths == halo widths */
    int xs, ys, zs;       /* source point */
    int x3, x4, y3, y4, z3, z4; /* inner3D range (whole grid without PML) */
} orc_fd_grid;

/* Index helpers following FDTDmacros.hpp:13-14 (k fastest, i slowest). */
static inline int64_t orc_idx3(const orc_fd_grid *g, int i, int j, int k)
{
    return (int64_t)g->nz * g->ny * i + (int64_t)g->nz * j + k;
}
static inline int64_t orc_idx3_l(const orc_fd_grid *g, int i, int j, int k)
{
    return (int64_t)(g->nz + 2 * g->lz) * (g->ny + 2 * g->ly) * (i + g->lx) +
           (int64_t)(g->nz + 2 * g->lz) * (j + g->ly) + (k + g->lz);
}

/* FDTDkernels.hpp:121-128 addRHS: pn[src] += vp[src]*rhs */
void orc_fd_add_rhs(const orc_fd_grid *g, float *pn, const float *vp, float rhs);
/* FDTDkernels.hpp:8-34 inner3D: p_ca <- 2 p_cb - p_ca + vp (coef0 p_cb + lap) */
void orc_fd_inner3d(const orc_fd_grid *g, const float *cx, const float *cy,
                    const float *cz, double coef0, const float *vp,
                    float *p_ca, const float *p_cb);
/* FDTDkernels.hpp:108-118 applySponge: p_ca *= s, p_cb *= s */
void orc_fd_apply_sponge(const orc_fd_grid *g, const float *sponge, float *p_ca,
                         float *p_cb);

/* The FD time loop (FDTDmain.cpp:55-81, output FDTDutils.hpp:193-210), run for
 * nsteps from step it0.  p0 is buffer i1 (previous), p1 buffer i2 (current) on
 * entry; on exit the roles have been swapped nsteps times, *cur_is_p1 says
 * which array holds the current field.  trace_src[s] receives the value printed
 * by output() after step it0+s; rec_val[r*nsteps+s] the value at receiver r.
 * Receivers are (i,j,k) interior coordinates in rec_ijk[3r..3r+2]. */
void orc_fd_run(const orc_fd_grid *g, const float *cx, const float *cy,
                const float *cz, double coef0, const float *vp,
                const float *sponge, const float *rhs, int it0, int nsteps,
                float *p0, float *p1, int *cur_is_p1, float *trace_src,
                int nrec, const int *rec_ijk, float *rec_val);

/* All-FP64 restatement of the same loop (fields, coefficients and the source
 * term in double).  Coefficients are the exact rationals / dx^2. */
void orc_fd_run_f64(const orc_fd_grid *g, const double *cx, const double *cy,
                    const double *cz, double coef0, const double *vp,
                    const double *sponge, const double *rhs, int it0, int nsteps,
                    double *p0, double *p1, int *cur_is_p1, double *trace_src,
                    int nrec, const int *rec_ijk, double *rec_val);

/* PML (FDTDkernels.hpp:36-104, FDTDutils.hpp:84-171).  pml_profile_init of
 * one axis (profile[0..n+1], grid -1..n); init_eta's scale; one
 * computeOneStepPML (pml3D on the slabs of width nd[3], inner3D inside, eta =
 * ex[i+1] + ey[j+1] + ez[k+1], hd2 = 1/(4 dx^2) per axis, phi interior-sized);
 * and the time loop with it. */
void orc_fd_pml_profile(int n, int nd, float scale, float *profile);
float orc_fd_pml_scale(float dt, float vmax, int nd, float d);
void orc_fd_pml_step(const orc_fd_grid *g, const int *nd, const float *cx, const float *cy, const float *cz,
                     double coef0, const float *vp, const float *ex, const float *ey, const float *ez,
                     const float *hd2, float *phi, float *p_ca, const float *p_cb);
void orc_fd_run_pml(const orc_fd_grid *g, const int *nd, const float *cx, const float *cy, const float *cz,
                    double coef0, const float *vp, const float *ex, const float *ey, const float *ez,
                    const float *hd2, float *phi, const float *rhs, int it0, int nsteps, float *p0, float *p1,
                    int *cur_is_p1, float *trace_src, int nrec, const int *rec_ijk, float *rec_val);

/* ------------------------------------------------------------------- SEM */
/* GLL points of order p (SEMQkGLBasisFunctions.hpp:36-78), from the roots of
 * P'_p by Newton iteration in double. */
void orc_gll_points(int p, double *xi);
/* GLL weights (SEMQkGLBasisFunctions.hpp:80-117).  ref_rounding=1 rounds the
 * exact weights 2/(p(p+1)P_p(xi)^2) to 8 decimals, which is the table the
 * reference carries; 0 returns the exact weights. */
void orc_gll_weights(int p, double *w, int ref_rounding);
/* dphi[j*(p+1)+i] = l'_j(xi_i), the derivative of the j-th Lagrange basis
 * polynomial at node i (SEMQkGLBasisFunctions.hpp:178-324 layout
 * derivativeBasisFunction1D(j,i)). */
void orc_gll_deriv(int p, const double *xi, double *dphi);

typedef struct {
    int ex, ey, ez;       /* elements per axis (SEMmesh.cpp:3-26) */
    float lx, ly, lz;     /* domain lengths */
    int order;
    int nx, ny, nz;       /* nodes per axis */
    float hx, hy, hz;     /* element sizes (float members of SEMmesh) */
} orc_sem_mesh;

void orc_sem_mesh_init(orc_sem_mesh *m, int ex, int ey, int ez, float lx,
                       float ly, float lz, int order);
/* SEMmesh.cpp:81-139 getCoordInOneDirection for element n (h passed as int) */
void orc_sem_coord_1d(int order, int h, int n, float *coord);
/* SEMmesh.cpp:205-231 globalNodesList: list[e*(p+1)^3 + dofLocal] */
void orc_sem_nodes_list(const orc_sem_mesh *m, int *list);
/* SEMmesh.cpp:266-285 getElementNumberFromPoints */
int orc_sem_element_from_point(const orc_sem_mesh *m, float x, float y, float z);
/* SEMmesh.cpp:314-344 getListOfInteriorNodes (3D); returns the count */
int orc_sem_interior_nodes(const orc_sem_mesh *m, int *list);

/* Reference-semantics SEM state: per-element FP32 B (6 Voigt entries per GLL
 * point) and local mass, computed as computeB V2 does
 * (SEMQkGLIntegrals.hpp:206-359) from the float element coordinates
 * (SEMmesh.cpp:168-202).  The reference recomputes them every step; they do
 * not change between steps, so they are computed once here. */
typedef struct orc_sem orc_sem;
orc_sem *orc_sem_create(const orc_sem_mesh *m, const float *model_per_element,
                        float dt);
void orc_sem_destroy(orc_sem *s);
int orc_sem_num_nodes(const orc_sem *s);
/* SEMsolver.cpp:21-125 computeOneStep (3D, sequential element order):
 * p_i1 <- 2 p_i2 - p_i1 - dt^2 y / M on interior nodes, after injecting
 * dt^2 c_e^2 rhs into p_i2 at the first node of the source element. */
void orc_sem_step(orc_sem *s, float *p_i1, float *p_i2, int src_element,
                  float rhs);
/* Diagnostics of the last step (assembled FP32 M and y). */
void orc_sem_get_my(const orc_sem *s, float *M, float *y);
/* threads > 1: orc_sem_step runs the element loop colour by colour on that
 * many OpenMP threads (no race; sums in colour order: for timing). */
void orc_sem_set_threads(orc_sem *s, int threads);
/* y = K p with the element loop of computeOneStep (no source, no update). */
void orc_sem_apply_k(orc_sem *s, const float *p, float *y, float *M);

/* All-FP64 SEM: exact GLL weights, exact box geometry, double fields,
 * sum-factorised element operator; same algorithm otherwise. */
typedef struct orc_sem64 orc_sem64;
orc_sem64 *orc_sem64_create(const orc_sem_mesh *m,
                            const double *model_per_element, double dt);
void orc_sem64_destroy(orc_sem64 *s);
void orc_sem64_step(orc_sem64 *s, double *p_i1, double *p_i2, int src_element,
                    double rhs);
void orc_sem64_apply_k(orc_sem64 *s, const double *p, double *y, double *M);

/* Dense FP64 element stiffness K_e (npts x npts, row-major) of a general hex
 * element given its (p+1)^3 node coordinates, by direct quadrature of
 * grad(phi_i).grad(phi_j) -- the brute-force check of the operator. */
void orc_sem_dense_element_f64(int p, const double *X, const double *Y,
                               const double *Z, double *K, double *Mloc);

/* ---------------------------------------------------------------- SEM2D */
/* The reference's SEM2D build (sem2d_oracle.c): ex x ez quadrilaterals of
 * order p on the x-z plane, node = ix + iz nx, element e = i + k ex, first-
 * order paraxial boundary on all four edges (SEMsolver.cpp:91-123). */
typedef struct orc_sem2d orc_sem2d;
orc_sem2d *orc_sem2d_create(int ex, int ez, float lx, float lz, int order, const float *model, float dt);
void orc_sem2d_destroy(orc_sem2d *s);
/* sizes[7] = nx, nz, nodes, elements, interior nodes, boundary nodes, faces */
void orc_sem2d_sizes(const orc_sem2d *s, int *sizes);
/* element node lists [nE][q*q], interior nodes, boundary nodes (in the
 * reference's order), faceInfos [nF][2+q], face node -> boundary index [nF][q];
 * any pointer may be NULL */
void orc_sem2d_lists(const orc_sem2d *s, int *nodes, int *interior, int *bnodes, int *face, int *fmap);
void orc_sem2d_coords(const orc_sem2d *s, float *cx, float *cz);
/* SEMsolver::computeOneStep with SEM2D: source, element loop, interior and
 * boundary updates of p_i1 from p_i2 */
void orc_sem2d_step(orc_sem2d *s, float *p_i1, float *p_i2, int src_element, float rhs);
void orc_sem2d_apply_k(orc_sem2d *s, const float *p, float *y, float *M);
/* the last step's M, y and boundary term Sh (per boundary node); NULL skips */
void orc_sem2d_get_my(const orc_sem2d *s, float *M, float *y, float *Sh);
void orc_sem2d_boundary_term(orc_sem2d *s, float *Sh);
/* 0 leaves the boundary term out: the edges reflect (a check of what the
 * paraxial term absorbs) */
void orc_sem2d_set_paraxial(orc_sem2d *s, int on);
/* SEMmesh.cpp:380-429: the reference's 4-colour element lists (2D); returns
 * the per-colour capacity, lists[c * capacity + k] */
int orc_sem2d_colors(int ex, int ez, int *counts, int *lists);
/* The 2D step in FP64 on exact box geometry with the same weights: the
 * yardstick of the FP32 paths' rounding. */
typedef struct orc_sem2d64 orc_sem2d64;
orc_sem2d64 *orc_sem2d64_create(int ex, int ez, float lx, float lz, int order, const double *model, double dt,
                                int src_element);
void orc_sem2d64_destroy(orc_sem2d64 *s);
void orc_sem2d64_step(orc_sem2d64 *s, double *p_i1, double *p_i2, double rhs);

#ifdef __cplusplus
}
#endif
#endif
