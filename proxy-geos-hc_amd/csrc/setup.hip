// setup.hip -- host-side setup of the hot path (coefficients, time step,
// source wavelet, velocity models).  Host code only; FP32/FP64 exactly as the
// reference types its expressions (compiled without FMA contraction).
//
//   init_coef        src/fd/utilsFDTD/FDTDutils.hpp:13-58
//   coef0            src/fd/utilsFDTD/FDTDinit.hpp:212-227
//   compute_dt_sch   src/fd/utilsFDTD/FDTDutils.hpp:59-76
//   evaluateRicker   src/utils/utils.hpp:11-46, computeSourceTerm 48-57
//   init_models      src/fd/utilsFDTD/FDTDinit.hpp:249-317
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "phc_common.hpp"
#include "phc_setup.h"

namespace {

// Central second-difference weights num/den for half widths 1..6, in the
// reference's operand types (float quotients; the two last entries of L=3 are
// double quotients; L=6 ends with the reference's 1/1663).
struct Frac { float num, den; bool dbl; };
const Frac kCoef[7][7] = {
    {},
    {{-2, 1, false}, {1, 1, false}},
    {{-5, 2, false}, {4, 3, false}, {-1, 12, false}},
    {{-49, 18, false}, {3, 2, false}, {-3, 20, true}, {1, 90, true}},
    {{-205, 72, false}, {8, 5, false}, {-1, 5, false}, {8, 315, false}, {-1, 560, false}},
    {{-5269, 1800, false}, {5, 3, false}, {-5, 21, false}, {5, 126, false}, {-5, 1008, false},
     {1, 3150, false}},
    {{-5369, 1800, false}, {12, 7, false}, {-15, 56, false}, {10, 189, false}, {-1, 112, false},
     {2, 1925, false}, {1, 1663, false}},
};

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

} // namespace

extern "C" {

phc_status phc_fd_init_coef(int L, float dx, float *coef)
{
    if (!coef) return phc::fail(PHC_ERR_INVALID, "coef is NULL");
    if (L < 1 || L > 6) return phc::fail(PHC_ERR_INVALID, "half width %d outside 1..6", L);
    const float dx2 = dx * dx;
    for (int l = 0; l <= L; ++l) {
        const Frac &f = kCoef[L][l];
        if (f.dbl) coef[l] = (float)((double)f.num / (double)f.den / (double)dx2);
        else coef[l] = f.num / f.den / dx2;
    }
    return PHC_OK;
}

double phc_fd_coef0(const float *cx, int Lx, const float *cy, int Ly, const float *cz, int Lz)
{
    float tx = 0.f, ty = 0.f, tz = 0.f;
    for (int i = 1; i <= Lx; ++i) tx += cx[i];
    for (int i = 1; i <= Ly; ++i) ty += cy[i];
    for (int i = 1; i <= Lz; ++i) tz += cz[i];
    return -2. * (double)((tx + ty) + tz);
}

float phc_fd_compute_dt(float vmax, const float *cx, int Lx, const float *cy, int Ly,
                        const float *cz, int Lz, float cfl)
{
    float ftmp = 0.f;
    ftmp += (std::fabs(cx[0]) + std::fabs(cy[0])) + std::fabs(cz[0]);
    for (int i = 1; i <= Lx; ++i) ftmp += 2.f * std::fabs(cx[i]);
    for (int i = 1; i <= Ly; ++i) ftmp += 2.f * std::fabs(cy[i]);
    for (int i = 1; i <= Lz; ++i) ftmp += 2.f * std::fabs(cz[i]);
    return (2.f * cfl) / (std::sqrt(ftmp) * vmax);
}

float phc_ricker(float time_n, float f0, int order, int exp_variant)
{
    const float o_tpeak = (float)(1.0 / (double)f0);
    if ((double)time_n <= -0.9 * (double)o_tpeak || (double)time_n >= 2.9 * (double)o_tpeak)
        return 0.f;
    const float pi = (float)M_PI;
    const float lam = (f0 * pi) * (f0 * pi);
    const float d = time_n - o_tpeak;
    const bool dexp = exp_variant == PHC_EXP_DOUBLE;
    switch (order) {
    case 2: {
        const float arg = ((-lam) * d) * d;
        const double e = dexp ? std::exp((double)arg) : (double)std::exp(arg);
        return (float)((2.0 * (double)lam) * (((2.0 * (double)lam) * (double)d) * (double)d - 1.0) * e);
    }
    case 1: {
        const float arg = ((-lam) * d) * d;
        const double e = dexp ? std::exp((double)arg) : (double)std::exp(arg);
        return (float)(((-2.0) * (double)lam) * (double)d * e);
    }
    case 0: {
        const float arg = ((-2.f * lam) * d) * d;
        if (dexp) return (float)((double)(-d) * std::exp((double)arg));
        return (-d) * std::exp(arg);
    }
    default:
        return 0.f;
    }
}

phc_status phc_source_term(int nsamples, float dt, float f0, int order, int exp_variant, float *out)
{
    if (nsamples < 0 || (nsamples > 0 && !out)) return phc::fail(PHC_ERR_INVALID, "bad output");
    if (order < 0 || order > 2) return phc::fail(PHC_ERR_INVALID, "Ricker order %d not 0..2", order);
    for (int i = 0; i < nsamples; ++i) out[i] = phc_ricker((float)i * dt, f0, order, exp_variant);
    return PHC_OK;
}

phc_status phc_fd_model_two_layer(int nx, int ny, int nz, float vmin, float dt, float *vp)
{
    if (!vp || nx < 1 || ny < 1 || nz < 1) return phc::fail(PHC_ERR_INVALID, "bad model arguments");
    const float v0 = ((vmin * vmin) * dt) * dt;
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t k = 0; k < nz; ++k)
                vp[(i * ny + j) * nz + k] = (k < nz / 2) ? v0 : 2 * v0;
    return PHC_OK;
}

phc_status phc_fd_model_layered(int nx, int ny, int nz, int nlayers, float vtop, float vbottom,
                                uint64_t seed, float dt, int *interfaces, float *vp)
{
    if (!vp || nx < 1 || ny < 1 || nz < 1 || nlayers < 1 || nlayers > nz)
        return phc::fail(PHC_ERR_INVALID, "bad layered-model arguments");
    // nlayers-1 distinct interface depths drawn uniformly from 1..nz-1,
    // sorted; layer velocities evenly spaced from vtop to vbottom.
    std::vector<int> itf;
    uint64_t s = seed;
    while ((int)itf.size() < nlayers - 1) {
        const int d = 1 + (int)(splitmix64(s) % (uint64_t)(nz - 1));
        bool dup = false;
        for (int x : itf) dup |= (x == d);
        if (!dup) itf.push_back(d);
    }
    for (size_t a = 0; a < itf.size(); ++a)
        for (size_t b = a + 1; b < itf.size(); ++b)
            if (itf[b] < itf[a]) std::swap(itf[a], itf[b]);
    if (interfaces)
        for (size_t a = 0; a < itf.size(); ++a) interfaces[a] = itf[a];
    std::vector<float> v2(nz);
    for (int k = 0; k < nz; ++k) {
        int layer = 0;
        while (layer < (int)itf.size() && k >= itf[layer]) ++layer;
        const float v = (nlayers == 1) ? vtop
                                       : (float)(vtop + (double)(vbottom - vtop) * layer / (nlayers - 1));
        v2[k] = ((v * v) * dt) * dt;
    }
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t k = 0; k < nz; ++k) vp[(i * ny + j) * nz + k] = v2[k];
    return PHC_OK;
}

phc_status phc_sem_model_gaussian(int ex, int ey, int ez, double sigma, float vmin, float vmax,
                                  uint64_t seed, float *model)
{
    if (!model || ex < 1 || ey < 1 || ez < 1 || !(sigma >= 0) || !(vmax >= vmin))
        return phc::fail(PHC_ERR_INVALID, "bad Gaussian-model arguments");
    const int64_t n = (int64_t)ex * ey * ez;
    std::vector<double> a(n), b(n);
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i += 2) {
        // Box-Muller on two uniforms in (0, 1]
        const double u1 = ((splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const double u2 = ((splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const double r = std::sqrt(-2.0 * std::log(u1));
        a[i] = r * std::cos(2.0 * M_PI * u2);
        if (i + 1 < n) a[i + 1] = r * std::sin(2.0 * M_PI * u2);
    }
    if (sigma > 0) {
        const int rad = (int)std::ceil(3.0 * sigma);
        std::vector<double> k(2 * rad + 1);
        double ks = 0;
        for (int d = -rad; d <= rad; ++d) ks += (k[d + rad] = std::exp(-0.5 * d * d / (sigma * sigma)));
        for (double &v : k) v /= ks;
        const int dims[3] = {ex, ez, ey};
        const int64_t strides[3] = {1, ex, (int64_t)ex * ez};
        for (int ax = 0; ax < 3; ++ax) {
            const int m = dims[ax];
            const int64_t st = strides[ax];
            for (int64_t i = 0; i < n; ++i) {
                const int c = (int)((i / st) % m);
                double acc = 0;
                for (int d = -rad; d <= rad; ++d) {
                    const int cc = std::min(m - 1, std::max(0, c + d));
                    acc += k[d + rad] * a[i + (int64_t)(cc - c) * st];
                }
                b[i] = acc;
            }
            std::swap(a, b);
        }
    }
    double lo = a[0], hi = a[0];
    for (double v : a) { lo = std::min(lo, v); hi = std::max(hi, v); }
    for (int64_t i = 0; i < n; ++i)
        model[i] = (hi > lo) ? (float)(vmin + (vmax - vmin) * (a[i] - lo) / (hi - lo)) : vmin;
    return PHC_OK;
}

// PML setup (FDTDinit.hpp:155-173, FDTDutils.hpp:84-107, 129-160), float as
// the reference computes it.
int phc_fd_pml_width(float vmin, float f0, float d)
{
    const float fmax = 2.5f * f0;
    const float lambdamax = vmin / fmax;
    return (int)(4 * lambdamax / d);  // ntaper = 4
}

float phc_fd_pml_scale(float dt, float vmax, int ndamp, float d)
{
    return dt * 3.f * vmax * std::log(1000.f) / (2.f * ndamp * d);
}

phc_status phc_fd_pml_profile(int n, int ndamp, float scale, float *profile)
{
    if (!profile) return phc::fail(PHC_ERR_INVALID, "profile is NULL");
    if (n < 1 || ndamp < 0 || 2 * ndamp > n + 2)
        return phc::fail(PHC_ERR_INVALID, "bad PML width %d for %d points", ndamp, n);
    // pml_profile_init(profile, 0, n+1, ndamp, ndamp, scale): quadratic ramps,
    // scale at the outermost index of each end
    for (int i = 0; i <= n + 1; ++i) profile[i] = 0.f;
    if (ndamp == 0) return PHC_OK;
    const float tmp = scale / (float)(ndamp * ndamp);
    for (int i = 1; i <= ndamp; ++i) profile[ndamp - i] = (float)(i * i) * tmp;
    for (int i = 1; i <= ndamp; ++i) profile[n + 2 - ndamp + i - 1] = (float)(i * i) * tmp;
    return PHC_OK;
}

int phc_sem_color_count(int ex, int ey, int ez, int *max_per_color)
{
    // SEMmesh.cpp:380-381
    if (ex < 1 || ez < 1 || ey < 0) return 0;
    const int m = (ex / 2 + ex % 2) * (ez / 2 + ez % 2) * (ey == 0 ? 1 : ey / 2 + ey % 2);
    if (max_per_color) *max_per_color = m;
    return ey == 0 ? 4 : 8;
}

phc_status phc_sem_color_elements(int ex, int ey, int ez, int *counts, int *lists)
{
    int mpc = 0;
    const int nc = phc_sem_color_count(ex, ey, ez, &mpc);
    if (nc == 0) return phc::fail(PHC_ERR_INVALID, "mesh %d x %d x %d has no elements", ex, ey, ez);
    const int ny = ey == 0 ? 1 : ey;
    for (int c = 0; c < nc; ++c) {
        // colour bits: x parity, z parity, y parity (2D: red, green, blue, yellow)
        const int px = c & 1, pz = (c >> 1) & 1, py = (c >> 2) & 1;
        int k = 0;
        for (int j = py; j < ny; j += 2)
            for (int l = pz; l < ez; l += 2)
                for (int i = px; i < ex; i += 2) {
                    if (lists) lists[(int64_t)c * mpc + k] = i + l * ex + j * ex * ez;
                    ++k;
                }
        if (counts) counts[c] = k;
    }
    return PHC_OK;
}

} // extern "C"
