// sem_setup.hpp -- one-time FP64 setup of the SEM hot path: GLL basis,
// structured box mesh, assembled 1D operators, lumped mass and the per-node
// factors the step kernels read.  Host code.
//
//   GLL points / weights / derivatives  src/sem/finiteElements/SEMQkGLBasisFunctions.hpp:36-324
//   mesh (node numbering, coordinates)  src/sem/mesh/SEMmesh.cpp:3-344
//   element Jacobian, B, local mass     src/sem/finiteElements/SEMQkGLIntegrals.hpp:206-359
//   global mass (sum of local / c^2)    src/sem/solver/SEMsolver.cpp:74-80
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace phc {

struct GllBasis {
    int p = 0;
    std::vector<double> xi, w, D;  // D[a*(p+1)+b] = l_a'(xi_b)
};

inline void legendre_pd(int p, double x, double &P, double &dP)
{
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    if (p == 0) { P = 1.0; dP = 0.0; return; }
    for (int n = 2; n <= p; ++n) {
        const double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
        const double d2 = d0 + (2.0 * n - 1.0) * p1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    }
    P = p1;
    dP = d1;
}

// GLL nodes = endpoints and roots of P_p'; weights 2/(p(p+1)P_p(x)^2),
// optionally rounded to 8 decimals as the reference's table carries them;
// D from the Lagrange-derivative formula at GLL nodes.
inline GllBasis make_gll(int p, bool ref_weights)
{
    GllBasis g;
    g.p = p;
    const int q = p + 1;
    g.xi.assign(q, 0.0);
    g.w.assign(q, 0.0);
    g.D.assign(q * q, 0.0);
    g.xi[0] = -1.0;
    g.xi[p] = 1.0;
    for (int i = 1; i < p; ++i) {
        double x = -std::cos(M_PI * i / p);
        for (int it = 0; it < 100; ++it) {
            double P, dP;
            legendre_pd(p, x, P, dP);
            const double d2P = (2.0 * x * dP - p * (p + 1.0) * P) / (1.0 - x * x);
            const double dx = dP / d2P;
            x -= dx;
            if (std::fabs(dx) < 1e-17) break;
        }
        g.xi[i] = x;
    }
    for (int i = 0; i <= p / 2; ++i) {
        const double a = 0.5 * (g.xi[p - i] - g.xi[i]);
        g.xi[i] = -a;
        g.xi[p - i] = a;
    }
    if (p % 2 == 0) g.xi[p / 2] = 0.0;
    std::vector<double> P(q);
    for (int i = 0; i < q; ++i) {
        double dP;
        legendre_pd(p, g.xi[i], P[i], dP);
        g.w[i] = 2.0 / (p * (p + 1.0) * P[i] * P[i]);
        if (ref_weights) g.w[i] = std::floor(g.w[i] * 1e8 + 0.5) / 1e8;
    }
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b) {
            double v;
            if (a != b) v = (P[b] / P[a]) / (g.xi[b] - g.xi[a]);
            else if (b == 0) v = -p * (p + 1.0) / 4.0;
            else if (b == p) v = p * (p + 1.0) / 4.0;
            else v = 0.0;
            g.D[a * q + b] = v;
        }
    return g;
}

// One axis of the structured box: ne elements of width h (the reference passes
// the element size to its coordinate routine as an int, SEMmesh.hpp:61), order
// p.  A[i][d+p] = (assembled 1D stiffness scaled by 2/h)[i][i+d] / Wg[i] and
// Wg[i] = assembled 1D mass scaled by h/2.
struct SemAxis {
    int ne = 0, n = 0;
    double h = 0.0, half = 0.0;
    std::vector<double> Wg;         // [n]
    std::vector<double> A;          // [n][2p+1]
};

inline SemAxis make_axis(const GllBasis &g, int ne, double h)
{
    const int p = g.p, q = p + 1, w2 = 2 * p + 1;
    SemAxis ax;
    ax.ne = ne;
    ax.n = ne * p + 1;
    ax.h = h;
    ax.half = h / 2.0;
    ax.Wg.assign(ax.n, 0.0);
    std::vector<double> Ag((size_t)ax.n * w2, 0.0);  // banded, unscaled by Wg
    std::vector<double> A1(q * q, 0.0);
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b) {
            double s = 0.0;
            for (int l = 0; l < q; ++l) s += g.w[l] * g.D[a * q + l] * g.D[b * q + l];
            A1[a * q + b] = s;
        }
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < q; ++a) {
            const int i = e * p + a;
            ax.Wg[i] += ax.half * g.w[a];
            for (int b = 0; b < q; ++b) Ag[(size_t)i * w2 + (b - a + p)] += A1[a * q + b] / ax.half;
        }
    ax.A.assign((size_t)ax.n * w2, 0.0);
    for (int i = 0; i < ax.n; ++i)
        for (int d = 0; d < w2; ++d) ax.A[(size_t)i * w2 + d] = Ag[(size_t)i * w2 + d] / ax.Wg[i];
    return ax;
}

} // namespace phc
