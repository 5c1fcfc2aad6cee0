/*
 * phc_setup.h -- host-side setup entry points of libphc.so (no device work).
 * They compute what the reference's FDTDInit/FDTDUtils/SolverUtils compute
 * before the time loop, with the reference's own operand types.
 */
#ifndef PHC_SETUP_H
#define PHC_SETUP_H

#include <stdint.h>

#include "phc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* exp variant of evaluateRicker: the FD translation unit binds ::exp(double),
 * the SEM one std::exp(float) (src/utils/utils.hpp:31-44 under each unit's
 * using-directives). */
#define PHC_EXP_DOUBLE 0
#define PHC_EXP_FLOAT 1

/* FDTDUtils::init_coef (src/fd/utilsFDTD/FDTDutils.hpp:13-58): coef[0..L]. */
phc_status phc_fd_init_coef(int L, float dx, float *coef);
/* FDTDInit::init_coefficients coef0 (src/fd/utilsFDTD/FDTDinit.hpp:212-227). */
double phc_fd_coef0(const float *cx, int Lx, const float *cy, int Ly, const float *cz, int Lz);
/* FDTDUtils::compute_dt_sch (src/fd/utilsFDTD/FDTDutils.hpp:59-76); the
 * reference fixes cfl = 0.8. */
float phc_fd_compute_dt(float vmax, const float *cx, int Lx, const float *cy, int Ly,
                        const float *cz, int Lz, float cfl);
/* SolverUtils::evaluateRicker (src/utils/utils.hpp:11-46). */
float phc_ricker(float time_n, float f0, int order, int exp_variant);
/* SolverUtils::computeSourceTerm (src/utils/utils.hpp:48-57). */
phc_status phc_source_term(int nsamples, float dt, float f0, int order, int exp_variant,
                           float *out);
/* FDTDInit::init_models (src/fd/utilsFDTD/FDTDinit.hpp:277-301): vp^2 dt^2,
 * doubled for k >= nz/2.  IDX3 order. */
phc_status phc_fd_model_two_layer(int nx, int ny, int nz, float vmin, float dt, float *vp);
/* Seeded layered model (no reference counterpart; the FD benchmark model):
 * nlayers layers along k with nlayers-1 distinct sorted interface depths drawn
 * with splitmix64(seed) from 1..nz-1, velocities evenly spaced vtop..vbottom.
 * Writes vp^2 dt^2 in IDX3 order; interfaces (may be NULL) gets the depths. */
phc_status phc_fd_model_layered(int nx, int ny, int nz, int nlayers, float vtop, float vbottom,
                                uint64_t seed, float dt, int *interfaces, float *vp);

/* Seeded smooth random model per SEM element (no reference counterpart; the
 * model of the order-4 precision study): standard-normal white noise from
 * splitmix64(seed) (Box-Muller) on the ex*ey*ez element grid, smoothed by a
 * separable Gaussian of std sigma elements (truncated at 3 sigma, clamped
 * edges), min-max scaled to [vmin, vmax].  Element order e = ix + iz*ex +
 * iy*ex*ez. */
phc_status phc_sem_model_gaussian(int ex, int ey, int ez, double sigma, float vmin, float vmax,
                                  uint64_t seed, float *model);

/* Element colouring of SEMmesh::getNumberOfElementsByColor / sortElementsByColor
 * (SEMmesh.cpp:380-429).  ey == 0 (the SEM2D mesh): the reference's 4 colours,
 * red (even x, even z), green (odd, even), blue (even, odd), yellow (odd, odd),
 * each list z-major as the reference fills it.  ey > 0: 8 colours
 * c = px + 2 pz + 4 py by index parities (the reference's 4-colour scheme
 * ignores ey and leaves every element layer but the first uncoloured), each
 * list ordered y, z, x with x fastest.  No two elements of a colour share a
 * node.  Returns the number of colours; counts[c] and
 * lists[c * max_per_color + k] are filled when non-NULL; max_per_color is
 * getNumberOfElementsByColor's value. */
int phc_sem_color_count(int ex, int ey, int ez, int *max_per_color);
phc_status phc_sem_color_elements(int ex, int ey, int ez, int *counts, int *lists);

/* PML setup of FDTDInit -usePML (FDTDinit.hpp:155-173) and FDTDUtils::init_eta
 * (FDTDutils.hpp:84-107, 129-160), float as the reference computes it:
 * width = (int)(4 * vmin/(2.5 f0) / d); scale = dt*3*vmax*logf(1000)/(2*ndamp*d);
 * profile[0..n+1] (grid -1..n) = 0 except quadratic ramps of ndamp points
 * at both ends, scale at profile[0] and profile[n+1]. */
int phc_fd_pml_width(float vmin, float f0, float d);
float phc_fd_pml_scale(float dt, float vmax, int ndamp, float d);
phc_status phc_fd_pml_profile(int n, int ndamp, float scale, float *profile);

#ifdef __cplusplus
}
#endif
#endif
