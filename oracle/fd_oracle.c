/*
 * fd_oracle.c -- CPU restatement of the reference FD hot path and its setup.
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 *
 * Compiled with -ffp-contract=off: every float/double operation below rounds
 * exactly where the reference's C++ expression rounds on an x86-64 build
 * without FMA.  Comments give the reference file:line of each restated piece.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ Ricker */

/* utils.hpp:11-46.  The reference evaluates exp on a float argument; in the
 * FD translation unit that call binds to ::exp(double), in the SEM one (where
 * `using namespace std` precedes utils.hpp) to std::exp(float). */
float orc_ricker(float time_n, float f0, int order, int use_double_exp)
{
    const float o_tpeak = (float)(1.0 / (double)f0);
    float pulse = 0.0f;
    if (((double)time_n <= -0.9 * (double)o_tpeak) ||
        ((double)time_n >= 2.9 * (double)o_tpeak))
        return pulse;
    const float pi = (float)M_PI;
    const float lam = (f0 * pi) * (f0 * pi);
    const float d = time_n - o_tpeak;
    switch (order) {
    case 2: {
        const float arg = ((-lam) * d) * d;
        const double e = use_double_exp ? exp((double)arg) : (double)expf(arg);
        pulse = (float)((2.0 * (double)lam) *
                        (((2.0 * (double)lam) * (double)d) * (double)d - 1.0) * e);
        break;
    }
    case 1: {
        const float arg = ((-lam) * d) * d;
        const double e = use_double_exp ? exp((double)arg) : (double)expf(arg);
        pulse = (float)(((-2.0) * (double)lam) * (double)d * e);
        break;
    }
    case 0: {
        const float arg = (((float)(-2) * lam) * d) * d;
        if (use_double_exp)
            pulse = (float)((double)(-d) * exp((double)arg));
        else
            pulse = (-d) * expf(arg);
        break;
    }
    default:
        break;
    }
    return pulse;
}

void orc_source_term(int nsamples, float dt, float f0, int order,
                     int use_double_exp, float *out)
{
    /* utils.hpp:48-57: time_n = i*timeSample in float */
    for (int i = 0; i < nsamples; i++) {
        const float t = (float)i * dt;
        out[i] = orc_ricker(t, f0, order, use_double_exp);
    }
}

double orc_ricker_f64(double t, double f0, int order)
{
    const double tp = 1.0 / f0;
    if (t <= -0.9 * tp || t >= 2.9 * tp)
        return 0.0;
    const double lam = (f0 * M_PI) * (f0 * M_PI);
    const double d = t - tp;
    switch (order) {
    case 2: return 2.0 * lam * (2.0 * lam * d * d - 1.0) * exp(-lam * d * d);
    case 1: return -2.0 * lam * d * exp(-lam * d * d);
    case 0: return -d * exp(-2.0 * lam * d * d);
    default: return 0.0;
    }
}

/* ------------------------------------------------------------ coefficients */

/* FDTDutils.hpp:13-58.  Central second-difference weights c_l/dx^2 for half
 * width L.  The operand types follow the reference: float quotients, except
 * L=3 whose last two entries are double quotients.  L=6's last entry is the
 * reference's 1/1663 (the textbook weight is -1/16632). */
void orc_fd_init_coef(int L, float dx, float *coef)
{
    const float dx2 = dx * dx;
    switch (L) {
    case 1:
        coef[0] = -2.f / dx2;
        coef[1] = 1.f / dx2;
        break;
    case 2:
        coef[0] = -5.f / 2.f / dx2;
        coef[1] = 4.f / 3.f / dx2;
        coef[2] = -1.f / 12.f / dx2;
        break;
    case 3:
        coef[0] = -49.f / 18.f / dx2;
        coef[1] = 3.f / 2.f / dx2;
        coef[2] = (float)(-3. / 20. / (double)dx2);
        coef[3] = (float)(1. / 90. / (double)dx2);
        break;
    case 4:
        coef[0] = -205.f / 72.f / dx2;
        coef[1] = 8.f / 5.f / dx2;
        coef[2] = -1.f / 5.f / dx2;
        coef[3] = 8.f / 315.f / dx2;
        coef[4] = -1.f / 560.f / dx2;
        break;
    case 5:
        coef[0] = -5269.f / 1800.f / dx2;
        coef[1] = 5.f / 3.f / dx2;
        coef[2] = -5.f / 21.f / dx2;
        coef[3] = 5.f / 126.f / dx2;
        coef[4] = -5.f / 1008.f / dx2;
        coef[5] = 1.f / 3150.f / dx2;
        break;
    case 6:
        coef[0] = -5369.f / 1800.f / dx2;
        coef[1] = 12.f / 7.f / dx2;
        coef[2] = -15.f / 56.f / dx2;
        coef[3] = 10.f / 189.f / dx2;
        coef[4] = -1.f / 112.f / dx2;
        coef[5] = 2.f / 1925.f / dx2;
        coef[6] = 1.f / 1663.f / dx2;
        break;
    default:
        break;
    }
}

/* FDTDinit.hpp:212-227 */
double orc_fd_coef0(const float *cx, int Lx, const float *cy, int Ly,
                    const float *cz, int Lz)
{
    float tx = 0, ty = 0, tz = 0;
    for (int i = 1; i <= Lx; i++) tx += cx[i];
    for (int i = 1; i <= Ly; i++) ty += cy[i];
    for (int i = 1; i <= Lz; i++) tz += cz[i];
    return -2. * (double)((tx + ty) + tz);
}

/* FDTDutils.hpp:59-76 (cfl = 0.8 there) */
float orc_fd_compute_dt(float vmax, const float *cx, int Lx, const float *cy,
                        int Ly, const float *cz, int Lz, float cfl)
{
    float ftmp = 0.f;
    ftmp += (fabsf(cx[0]) + fabsf(cy[0])) + fabsf(cz[0]);
    for (int i = 1; i <= Lx; i++) ftmp += 2.f * fabsf(cx[i]);
    for (int i = 1; i <= Ly; i++) ftmp += 2.f * fabsf(cy[i]);
    for (int i = 1; i <= Lz; i++) ftmp += 2.f * fabsf(cz[i]);
    return (2.f * cfl) / (sqrtf(ftmp) * vmax);
}

/* ------------------------------------------------------------------ sponge */

void orc_fd_sponge_table(int L, float alpha, float *tab)
{
    /* FDTDinit.hpp:351: exp(alpha*pow((L-i),2)) -- alpha float promoted to
     * double, pow in double, exp in double, stored float. */
    for (int i = 0; i < L; i++)
        tab[i] = (float)exp((double)alpha * pow((double)(L - i), 2.0));
}

/* FDTDinit.hpp:319-396.  The loops are restated literally (including their
 * write order, which decides overlapping writes on grids narrower than 2L). */
void orc_fd_sponge(int nx, int ny, int nz, int L, float alpha, float *s)
{
#define S3(i, j, k) s[(int64_t)nz * ny * (i) + (int64_t)nz * (j) + (k)]
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) S3(i, j, k) = 1;
    for (int k = L; k < nz - L; k++)
        for (int j = L; j < ny - L; j++) {
            for (int i = 0; i < L; i++)
                S3(i, j, k) = (float)exp((double)alpha * pow((double)(L - i), 2.0));
            for (int i = nx - L; i < nx; i++)
                S3(i, j, k) = (float)exp((double)alpha * pow((double)(L - (nx - i)), 2.0));
        }
    for (int k = L; k < nz - L; k++)
        for (int i = L; i < nx - L; i++) {
            for (int j = 0; j < L; j++)
                S3(i, j, k) = (float)exp((double)alpha * pow((double)(L - j), 2.0));
            for (int j = ny - L; j < ny; j++)
                S3(i, j, k) = (float)exp((double)alpha * pow((double)(L - (ny - j)), 2.0));
        }
    for (int j = L; j < ny - L; j++)
        for (int i = L; i < nx - L; i++) {
            for (int k = 0; k < L; k++)
                S3(i, j, k) = (float)exp((double)alpha * pow((double)(L - k), 2.0));
            for (int k = nz - L; k < nz; k++)
                S3(i, j, k) = (float)exp((double)alpha * pow((double)(L - (nz - k)), 2.0));
        }
#undef S3
}

/* -------------------------------------------------------------------- step */

void orc_fd_add_rhs(const orc_fd_grid *g, float *pn, const float *vp, float rhs)
{
    /* FDTDkernels.hpp:121-128 */
    const int64_t a = orc_idx3_l(g, g->xs, g->ys, g->zs);
    const float v = vp[orc_idx3(g, g->xs, g->ys, g->zs)];
    pn[a] += v * rhs;
}

void orc_fd_inner3d(const orc_fd_grid *g, const float *cx, const float *cy,
                    const float *cz, double coef0, const float *vp,
                    float *p_ca, const float *p_cb)
{
    /* FDTDkernels.hpp:8-34.  lapx/lapy/lapz are float accumulators; the final
     * expression is evaluated in double (coef0 is double, 2. is a double
     * literal) and stored as float. */
    const int64_t sj = g->nz + 2 * g->lz;
    const int64_t si = sj * (g->ny + 2 * g->ly);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = g->x3; i < g->x4; i++)
        for (int j = g->y3; j < g->y4; j++)
            for (int k = g->z3; k < g->z4; k++) {
                const int64_t c = orc_idx3_l(g, i, j, k);
                float lapx = 0;
                for (int l = 1; l <= g->lx; l++)
                    lapx += cx[l] * (p_cb[c + l * si] + p_cb[c - l * si]);
                float lapy = 0;
                for (int l = 1; l <= g->ly; l++)
                    lapy += cy[l] * (p_cb[c + l * sj] + p_cb[c - l * sj]);
                float lapz = 0;
                for (int l = 1; l <= g->lz; l++)
                    lapz += cz[l] * (p_cb[c + l] + p_cb[c - l]);
                const double pc = (double)p_cb[c];
                p_ca[c] = (float)((2. * pc - (double)p_ca[c]) +
                                  (double)vp[orc_idx3(g, i, j, k)] *
                                      (((coef0 * pc + (double)lapx) + (double)lapy) +
                                       (double)lapz));
            }
}

void orc_fd_apply_sponge(const orc_fd_grid *g, const float *sponge, float *p_ca,
                         float *p_cb)
{
    /* FDTDkernels.hpp:108-118 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = g->x3; i < g->x4; i++)
        for (int j = g->y3; j < g->y4; j++)
            for (int k = g->z3; k < g->z4; k++) {
                const int64_t c = orc_idx3_l(g, i, j, k);
                const float s = sponge[orc_idx3(g, i, j, k)];
                p_ca[c] = p_ca[c] * s;
                p_cb[c] = p_cb[c] * s;
            }
}

void orc_fd_run(const orc_fd_grid *g, const float *cx, const float *cy,
                const float *cz, double coef0, const float *vp,
                const float *sponge, const float *rhs, int it0, int nsteps,
                float *p0, float *p1, int *cur_is_p1, float *trace_src,
                int nrec, const int *rec_ijk, float *rec_val)
{
    /* FDTDmain.cpp:55-81: addRHS(i2) -> computeOneStepSB(i1,i2) -> swap ->
     * output(i2).  prev/cur play the roles of pnGlobal(:,i1)/pnGlobal(:,i2). */
    float *prev = *cur_is_p1 ? p0 : p1;
    float *cur = *cur_is_p1 ? p1 : p0;
    for (int s = 0; s < nsteps; s++) {
        const int it = it0 + s;
        orc_fd_add_rhs(g, cur, vp, rhs[it]);
        orc_fd_inner3d(g, cx, cy, cz, coef0, vp, prev, cur);
        orc_fd_apply_sponge(g, sponge, prev, cur);
        float *t = prev; prev = cur; cur = t;
        if (trace_src)
            trace_src[s] = cur[orc_idx3_l(g, g->xs, g->ys, g->zs)];
        for (int r = 0; r < nrec; r++)
            rec_val[(int64_t)r * nsteps + s] =
                cur[orc_idx3_l(g, rec_ijk[3 * r], rec_ijk[3 * r + 1], rec_ijk[3 * r + 2])];
    }
    *cur_is_p1 = (cur == p1);
}

void orc_fd_run_f64(const orc_fd_grid *g, const double *cx, const double *cy,
                    const double *cz, double coef0, const double *vp,
                    const double *sponge, const double *rhs, int it0, int nsteps,
                    double *p0, double *p1, int *cur_is_p1, double *trace_src,
                    int nrec, const int *rec_ijk, double *rec_val)
{
    double *prev = *cur_is_p1 ? p0 : p1;
    double *cur = *cur_is_p1 ? p1 : p0;
    const int64_t sj = g->nz + 2 * g->lz;
    const int64_t si = sj * (g->ny + 2 * g->ly);
    for (int s = 0; s < nsteps; s++) {
        const int it = it0 + s;
        cur[orc_idx3_l(g, g->xs, g->ys, g->zs)] +=
            vp[orc_idx3(g, g->xs, g->ys, g->zs)] * rhs[it];
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = g->x3; i < g->x4; i++)
            for (int j = g->y3; j < g->y4; j++)
                for (int k = g->z3; k < g->z4; k++) {
                    const int64_t c = orc_idx3_l(g, i, j, k);
                    double lx = 0, ly = 0, lz = 0;
                    for (int l = 1; l <= g->lx; l++)
                        lx += cx[l] * (cur[c + l * si] + cur[c - l * si]);
                    for (int l = 1; l <= g->ly; l++)
                        ly += cy[l] * (cur[c + l * sj] + cur[c - l * sj]);
                    for (int l = 1; l <= g->lz; l++)
                        lz += cz[l] * (cur[c + l] + cur[c - l]);
                    prev[c] = 2. * cur[c] - prev[c] +
                              vp[orc_idx3(g, i, j, k)] * (coef0 * cur[c] + lx + ly + lz);
                }
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = g->x3; i < g->x4; i++)
            for (int j = g->y3; j < g->y4; j++)
                for (int k = g->z3; k < g->z4; k++) {
                    const int64_t c = orc_idx3_l(g, i, j, k);
                    const double sp = sponge[orc_idx3(g, i, j, k)];
                    prev[c] *= sp;
                    cur[c] *= sp;
                }
        double *t = prev; prev = cur; cur = t;
        if (trace_src)
            trace_src[s] = cur[orc_idx3_l(g, g->xs, g->ys, g->zs)];
        for (int r = 0; r < nrec; r++)
            rec_val[(int64_t)r * nsteps + s] =
                cur[orc_idx3_l(g, rec_ijk[3 * r], rec_ijk[3 * r + 1], rec_ijk[3 * r + 2])];
    }
    *cur_is_p1 = (cur == p1);
}

/* ------------------------------------------------------------------ PML */

void orc_fd_pml_profile(int n, int nd, float scale, float *profile)
{
    /* FDTDutils.hpp:84-107 pml_profile_init(profile, 0, n+1, nd, nd, scale):
     * indices 0..n+1 (grid -1..n); quadratic ramps of nd points at each end,
     * scale at the outermost index. */
    const int i_min = 0, i_max = n + 1;
    const int nn = i_max - i_min + 1, shift = i_min - 1;
    const int first_beg = 1 + shift, first_end = nd + shift;
    const int last_beg = nn - nd + 1 + shift, last_end = nn + shift;
    for (int i = i_min; i <= i_max; i++) profile[i] = 0.f;
    const int w = first_end - first_beg + 1;
    const float tmp = scale / (float)(w * w);
    for (int i = 1; i <= first_end - first_beg + 1; i++) profile[first_end - i + 1] = (float)(i * i) * tmp;
    for (int i = 1; i <= last_end - last_beg + 1; i++) profile[last_beg + i - 1] = (float)(i * i) * tmp;
}

float orc_fd_pml_scale(float dt, float vmax, int nd, float d)
{
    /* FDTDutils.hpp:149-160: dt_sch*3.f*vmax*logf(1000.f)/(2.f*ndamp*dx) in float */
    return dt * 3.f * vmax * logf(1000.f) / (2.f * nd * d);
}

/* eta at grid (i,j,k), i in -1..nx: the stored etax[i+1] + etay[j+1] + etaz[k+1]
 * (FDTDutils.hpp:110-127, 162-170: the extension writes the sum of the three
 * profiles on every region, and the profiles vanish on the inner box) */
static inline float pml_eta(const float *ex, const float *ey, const float *ez, int i, int j, int k)
{
    return ex[i + 1] + ey[j + 1] + ez[k + 1];
}

void orc_fd_pml_step(const orc_fd_grid *g, const int *nd, const float *cx, const float *cy, const float *cz,
                     double coef0, const float *vp, const float *ex, const float *ey, const float *ez,
                     const float *hd2, float *phi, float *p_ca, const float *p_cb)
{
    /* FDTDkernels.hpp:87-104 computeOneStepPML: pml3D on the six slabs of
     * width nd (FDTDinit.hpp:167-198 region bounds), inner3D on the rest.
     * Every point reads level cb and its own point of ca and phi, so the
     * order of the seven region loops does not matter. */
    const int64_t sj = g->nz + 2 * g->lz;
    const int64_t si = sj * (g->ny + 2 * g->ly);
    const int x3 = nd[0], x4 = g->nx - nd[0], y3 = nd[1], y4 = g->ny - nd[1], z3 = nd[2], z4 = g->nz - nd[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < g->nx; i++)
        for (int j = 0; j < g->ny; j++)
            for (int k = 0; k < g->nz; k++) {
                const int64_t c = orc_idx3_l(g, i, j, k);
                const int64_t q = orc_idx3(g, i, j, k);
                float lapx = 0;
                for (int l = 1; l <= g->lx; l++) lapx += cx[l] * (p_cb[c + l * si] + p_cb[c - l * si]);
                float lapy = 0;
                for (int l = 1; l <= g->ly; l++) lapy += cy[l] * (p_cb[c + l * sj] + p_cb[c - l * sj]);
                float lapz = 0;
                for (int l = 1; l <= g->lz; l++) lapz += cz[l] * (p_cb[c + l] + p_cb[c - l]);
                const double pc = (double)p_cb[c];
                if (i >= x3 && i < x4 && j >= y3 && j < y4 && k >= z3 && k < z4) {
                    /* inner3D (FDTDkernels.hpp:8-34) */
                    p_ca[c] = (float)((2. * pc - (double)p_ca[c]) +
                                      (double)vp[q] * (((coef0 * pc + (double)lapx) + (double)lapy) + (double)lapz));
                    continue;
                }
                /* pml3D (FDTDkernels.hpp:36-73) */
                const float lap = (float)(((coef0 * pc + (double)lapx) + (double)lapy) + (double)lapz);
                const float e = pml_eta(ex, ey, ez, i, j, k);
                const double num = (((2. - (double)(e * e)) + 2. * (double)e) * pc - (double)p_ca[c]) +
                                   (double)(vp[q] * (phi[q] + lap));
                p_ca[c] = (float)(num / (1. + 2. * (double)e));
                const float t = (pml_eta(ex, ey, ez, i + 1, j, k) - pml_eta(ex, ey, ez, i - 1, j, k)) *
                                    (p_cb[c + si] - p_cb[c - si]) * hd2[0] +
                                (pml_eta(ex, ey, ez, i, j + 1, k) - pml_eta(ex, ey, ez, i, j - 1, k)) *
                                    (p_cb[c + sj] - p_cb[c - sj]) * hd2[1] +
                                (pml_eta(ex, ey, ez, i, j, k + 1) - pml_eta(ex, ey, ez, i, j, k - 1)) *
                                    (p_cb[c + 1] - p_cb[c - 1]) * hd2[2];
                phi[q] = (float)((double)(phi[q] - t) / (1. + (double)e));
            }
}

void orc_fd_run_pml(const orc_fd_grid *g, const int *nd, const float *cx, const float *cy, const float *cz,
                    double coef0, const float *vp, const float *ex, const float *ey, const float *ez,
                    const float *hd2, float *phi, const float *rhs, int it0, int nsteps, float *p0, float *p1,
                    int *cur_is_p1, float *trace_src, int nrec, const int *rec_ijk, float *rec_val)
{
    /* FDTDmain.cpp:55-81 with usePML: addRHS(i2) -> computeOneStepPML(i1,i2) -> swap */
    float *prev = *cur_is_p1 ? p0 : p1;
    float *cur = *cur_is_p1 ? p1 : p0;
    for (int s = 0; s < nsteps; s++) {
        orc_fd_add_rhs(g, cur, vp, rhs[it0 + s]);
        orc_fd_pml_step(g, nd, cx, cy, cz, coef0, vp, ex, ey, ez, hd2, phi, prev, cur);
        float *t = prev; prev = cur; cur = t;
        if (trace_src) trace_src[s] = cur[orc_idx3_l(g, g->xs, g->ys, g->zs)];
        for (int r = 0; r < nrec; r++)
            rec_val[(int64_t)r * nsteps + s] =
                cur[orc_idx3_l(g, rec_ijk[3 * r], rec_ijk[3 * r + 1], rec_ijk[3 * r + 2])];
    }
    *cur_is_p1 = (cur == p1);
}
