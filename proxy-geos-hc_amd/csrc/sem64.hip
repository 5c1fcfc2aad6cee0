// sem64.hip -- FP64 SEM time loop (phc_sem64_* C ABI): the reference's SEM
// step (SEMsolver::computeOneStep, src/sem/solver/SEMsolver.cpp:21-125; the
// SEMproxy loop SEMproxy.cpp:41-119) on the structured box mesh with every
// field, mass and operator in double.  It is the accuracy yardstick of the
// FP32 paths (the order-4 precision study), not a throughput path.
//
// The operator is the separable form of the FP32 separable kernel (sem.hip):
// on the axis-aligned box the assembled stiffness is Ax Wz Wy + Wx Az Wy +
// Wx Wz Ay, so with Q = dt^2 Wx Wy Wz / M the update of an interior node is
//     pnew = 2 p - pold - Q (Lx p + Ly p + Lz p)
// with radius-p rows that depend on the index modulo p.  One thread per
// (ix, iz) column marches a chunk of iy with the iy-neighbours in registers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fd_kernels.hpp"
#include "phc_common.hpp"
#include "sem_setup.hpp"

namespace phc {

constexpr int S64_MAX_P = 5;
constexpr int S64_CHUNK = 16;

struct Sem64Args {
    const double *pn;
    double *pnm1;
    const double *qf;
    FdLayout L;  // i = iy, j = iz, k = ix; halo p
    double cx[S64_MAX_P][2 * S64_MAX_P + 1], cy[S64_MAX_P][2 * S64_MAX_P + 1], cz[S64_MAX_P][2 * S64_MAX_P + 1];
};

template <int P>
__global__ void __launch_bounds__(256) sem64_step_kernel(Sem64Args a)
{
    constexpr int W = 2 * P + 1;
    const FdLayout &L = a.L;
    const int k = blockIdx.x * 64 + threadIdx.x;  // ix
    const int j = blockIdx.y * 4 + threadIdx.y;   // iz
    const int i0 = blockIdx.z * S64_CHUNK;        // iy
    if (k < 1 || k >= L.nz - 1 || j < 1 || j >= L.ny - 1) return;  // boundary nodes stay 0
    const int i1 = min(i0 + S64_CHUNK, L.nx);
    double q[W];
#pragma unroll
    for (int l = 0; l < W - 1; ++l) q[l + 1] = a.pn[fd_off(L, i0 - P + l, j, k)];
    const int mx = k % P, mz = j % P;
    for (int i = i0; i < i1; ++i) {
#pragma unroll
        for (int l = 0; l < W - 1; ++l) q[l] = q[l + 1];
        q[W - 1] = a.pn[fd_off(L, i + P, j, k)];
        if (i < 1 || i >= L.nx - 1) continue;
        const int64_t c = fd_off(L, i, j, k);
        const int my = i % P;
        double lap = 0.0;
#pragma unroll
        for (int d = 0; d < W; ++d) {
            lap += a.cy[my][d] * q[d];
            lap += a.cz[mz][d] * a.pn[c + (int64_t)(d - P) * L.pitch_k];
            lap += a.cx[mx][d] * a.pn[c + (d - P)];
        }
        a.pnm1[c] = 2.0 * q[P] - a.pnm1[c] - a.qf[c] * lap;
    }
}

__global__ void sem64_inject_kernel(double *p, int64_t off, double coef, const double *rhs, int it)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[off] += coef * rhs[it];
}

__global__ void sem64_sample_kernel(const double *p, const int64_t *offs, int nrec, double *out, int stride,
                                    int step, int64_t src_off, double *src_trace)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) src_trace[step] = p[src_off];
    if (r < nrec) out[(int64_t)r * stride + step] = p[offs[r]];
}

using Sem64Fn = void (*)(Sem64Args);

static Sem64Fn sem64_pick(int p)
{
    switch (p) {
    case 1: return sem64_step_kernel<1>;
    case 2: return sem64_step_kernel<2>;
    case 3: return sem64_step_kernel<3>;
    case 4: return sem64_step_kernel<4>;
    case 5: return sem64_step_kernel<5>;
    default: return nullptr;
    }
}

} // namespace phc

using namespace phc;

struct phc_sem64 {
    int ex, ey, ez, p, nx, ny, nz;
    int64_t nn;
    int nsamples;
    FdLayout L;
    Sem64Args args;
    DevBuf lev[2], qf, rhs, trace, rec, recoff;
    int64_t src_off;
    double inj_coef;
    int i1 = 0, i2 = 1;
    int nrec = 0, last_n = 0, last_nrec = 0;
    hipStream_t stream = nullptr;
};

static int64_t s64_off(const phc_sem64 *h, int64_t node)
{
    const int ix = (int)(node % h->nx), iz = (int)((node / h->nx) % h->nz),
              iy = (int)(node / ((int64_t)h->nx * h->nz));
    return fd_off(h->L, iy, iz, ix);
}

extern "C" {

phc_status phc_sem64_create(const phc_sem_params *p, phc_sem64 **out)
{
    if (!out) return fail(PHC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!p) return fail(PHC_ERR_INVALID, "params is NULL");
    if (p->ex < 1 || p->ey < 1 || p->ez < 1) return fail(PHC_ERR_INVALID, "mesh has no elements");
    if (p->order < 1 || p->order > S64_MAX_P) return fail(PHC_ERR_INVALID, "order %d outside 1..%d", p->order, S64_MAX_P);
    if (!(p->lx > 0) || !(p->ly > 0) || !(p->lz > 0)) return fail(PHC_ERR_INVALID, "bad domain size");
    if (!p->model || p->nsamples < 0 || (p->nsamples > 0 && !p->rhs))
        return fail(PHC_ERR_INVALID, "model or rhs is NULL");
    const int64_t nE = (int64_t)p->ex * p->ey * p->ez;
    if (p->src_element < 0 || p->src_element >= nE) return fail(PHC_ERR_INVALID, "source element out of range");
    for (int64_t e = 0; e < nE; ++e)
        if (!(p->model[e] > 0.f) || !std::isfinite(p->model[e]))
            return fail(PHC_ERR_INVALID, "wave speed of element %lld is not positive", (long long)e);
    const int hxi = (int)(float)((double)p->lx / (1. * p->ex));
    const int hyi = (int)(float)((double)p->ly / (1. * p->ey));
    const int hzi = (int)(float)((double)p->lz / (1. * p->ez));
    if (hxi < 1 || hyi < 1 || hzi < 1) return fail(PHC_ERR_INVALID, "element size below 1");
    phc_sem64 *h = new phc_sem64();
    auto bail = [&](phc_status s) { phc_sem64_destroy(h); return s; };
    const int P = p->order, qn = P + 1, W = 2 * P + 1;
    h->ex = p->ex; h->ey = p->ey; h->ez = p->ez; h->p = P;
    h->nx = p->ex * P + 1;
    h->ny = p->ey * P + 1;
    h->nz = p->ez * P + 1;
    h->nn = (int64_t)h->nx * h->ny * h->nz;
    h->nsamples = p->nsamples;
    {
        const int e = p->src_element;
        const int exi = e % h->ex, ezi = (e / h->ex) % h->ez, eyi = e / (h->ex * h->ez);
        const int sx = exi * P, sz = ezi * P, sy = eyi * P;
        if (!(sx >= 1 && sx < h->nx - 1 && sz >= 1 && sz < h->nz - 1 && sy >= 1 && sy < h->ny - 1))
            return bail(fail(PHC_ERR_INVALID, "the FP64 path needs the source on an interior node"));
        const double c = p->model[e];
        const double dt = p->dt;
        h->inj_coef = dt * dt * c * c;  // SEMsolver.cpp:41-42, in double
        h->L = make_fd_layout(h->ny, h->nz, h->nx, P);
        h->src_off = fd_off(h->L, sy, sz, sx);
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "hipStreamCreate failed"));
    const GllBasis g = make_gll(P, p->weights_ref != 0);
    const SemAxis ax = make_axis(g, h->ex, hxi), ay = make_axis(g, h->ey, hyi), az = make_axis(g, h->ez, hzi);
    std::vector<double> M((size_t)h->nn, 0.0);
    for (int64_t e = 0; e < nE; ++e) {
        const int exi = (int)(e % h->ex), ezi = (int)((e / h->ex) % h->ez), eyi = (int)(e / ((int64_t)h->ex * h->ez));
        const double ic2 = 1.0 / ((double)p->model[e] * p->model[e]);
        for (int c = 0; c < qn; ++c)
            for (int b = 0; b < qn; ++b)
                for (int a = 0; a < qn; ++a) {
                    const int64_t node = (exi * P + a) + (int64_t)(ezi * P + b) * h->nx +
                                         (int64_t)(eyi * P + c) * h->nx * h->nz;
                    M[node] += (g.w[a] * ax.half) * (g.w[b] * az.half) * (g.w[c] * ay.half) * ic2;
                }
    }
    const FdLayout &L = h->L;
    const double dt2 = (double)p->dt * (double)p->dt;
    std::vector<double> buf((size_t)L.total, 0.0);
    phc_status st;
    const size_t bytes = sizeof(double) * (size_t)L.total;
    if ((st = h->lev[0].alloc(bytes)) || (st = h->lev[1].alloc(bytes)) || (st = h->qf.alloc(bytes)))
        return bail(st);
    if (hipMemcpy(h->lev[0].ptr, buf.data(), bytes, hipMemcpyHostToDevice) ||
        hipMemcpy(h->lev[1].ptr, buf.data(), bytes, hipMemcpyHostToDevice))
        return bail(fail(PHC_ERR_HIP, "upload failed"));
    for (int iy = 1; iy < h->ny - 1; ++iy)
        for (int iz = 1; iz < h->nz - 1; ++iz)
            for (int ix = 1; ix < h->nx - 1; ++ix) {
                const int64_t n = ix + (int64_t)iz * h->nx + (int64_t)iy * h->nx * h->nz;
                buf[(size_t)fd_off(L, iy, iz, ix)] = dt2 * ax.Wg[ix] * az.Wg[iz] * ay.Wg[iy] / M[n];
            }
    if (hipMemcpy(h->qf.ptr, buf.data(), bytes, hipMemcpyHostToDevice)) return bail(fail(PHC_ERR_HIP, "upload failed"));
    std::vector<double> rhs(p->rhs, p->rhs + p->nsamples);
    if ((st = h->rhs.alloc(sizeof(double) * std::max(1, p->nsamples)))) return bail(st);
    if (p->nsamples > 0 && hipMemcpy(h->rhs.ptr, rhs.data(), sizeof(double) * p->nsamples, hipMemcpyHostToDevice))
        return bail(fail(PHC_ERR_HIP, "upload failed"));
    Sem64Args &a = h->args;
    std::memset(&a, 0, sizeof(a));
    a.L = L;
    a.qf = h->qf.as<double>();
    auto rows = [&](const SemAxis &A, double (*c)[2 * S64_MAX_P + 1]) {
        for (int m = 0; m < P; ++m) {
            int i = m;
            if (i == 0) i = (A.n > P + 1) ? P : 0;
            if (i == 0 || i >= A.n - 1) continue;
            for (int d = 0; d < W; ++d) c[m][d] = A.A[(size_t)i * W + d];
        }
    };
    rows(ax, a.cx);
    rows(ay, a.cy);
    rows(az, a.cz);
    *out = h;
    return PHC_OK;
}

phc_status phc_sem64_destroy(phc_sem64 *h)
{
    if (!h) return PHC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (DevBuf *b : {&h->lev[0], &h->lev[1], &h->qf, &h->rhs, &h->trace, &h->rec, &h->recoff}) b->release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PHC_OK;
}

phc_status phc_sem64_set_receivers(phc_sem64 *h, int nrec, const int *nodes)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nrec < 0 || (nrec > 0 && !nodes)) return fail(PHC_ERR_INVALID, "bad receiver list");
    std::vector<int64_t> offs(std::max(nrec, 1), 0);
    for (int r = 0; r < nrec; ++r) {
        if (nodes[r] < 0 || nodes[r] >= h->nn) return fail(PHC_ERR_INVALID, "receiver node out of range");
        offs[r] = s64_off(h, nodes[r]);
    }
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_TRY(h->recoff.alloc(sizeof(int64_t) * offs.size()));
    PHC_HIP(hipMemcpy(h->recoff.ptr, offs.data(), sizeof(int64_t) * offs.size(), hipMemcpyHostToDevice));
    h->nrec = nrec;
    return PHC_OK;
}

// SEMproxy::run for samples it0 .. it0+nsteps-1: inject rhs[it] into level
// i2 at the source node, step (level i1 <- new), record, swap.
phc_status phc_sem64_run(phc_sem64 *h, int it0, int nsteps)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nsteps < 0 || it0 < 0 || it0 + nsteps > h->nsamples)
        return fail(PHC_ERR_INVALID, "steps %d..%d outside the %d samples", it0, it0 + nsteps - 1, h->nsamples);
    if (nsteps == 0) return PHC_OK;
    PHC_TRY(h->trace.alloc(sizeof(double) * nsteps));
    if (h->nrec > 0) PHC_TRY(h->rec.alloc(sizeof(double) * (size_t)nsteps * h->nrec));
    Sem64Fn fn = sem64_pick(h->p);
    if (!fn) return fail(PHC_ERR_INVALID, "no FP64 kernel for p=%d", h->p);
    const FdLayout &L = h->L;
    dim3 grid((L.nz + 63) / 64, (L.ny + 3) / 4, (L.nx + S64_CHUNK - 1) / S64_CHUNK);
    for (int s = 0; s < nsteps; ++s) {
        hipLaunchKernelGGL(sem64_inject_kernel, dim3(1), dim3(64), 0, h->stream, h->lev[h->i2].as<double>(),
                           h->src_off, h->inj_coef, h->rhs.as<double>(), it0 + s);
        Sem64Args a = h->args;
        a.pn = h->lev[h->i2].as<double>();
        a.pnm1 = h->lev[h->i1].as<double>();
        hipLaunchKernelGGL(fn, grid, dim3(64, 4), 0, h->stream, a);
        hipLaunchKernelGGL(sem64_sample_kernel, dim3(std::max(1, (h->nrec + 255) / 256)), dim3(256), 0, h->stream,
                           h->lev[h->i1].as<double>(), h->recoff.as<int64_t>(), h->nrec, h->rec.as<double>(), nsteps,
                           s, h->src_off, h->trace.as<double>());
        PHC_HIP(hipGetLastError());
        std::swap(h->i1, h->i2);
    }
    h->last_n = nsteps;
    h->last_nrec = h->nrec;
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

phc_status phc_sem64_get_traces(phc_sem64 *h, double *src, double *rec)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    if (src && h->last_n) PHC_HIP(hipMemcpy(src, h->trace.ptr, sizeof(double) * h->last_n, hipMemcpyDeviceToHost));
    if (rec && h->last_n && h->last_nrec)
        PHC_HIP(hipMemcpy(rec, h->rec.ptr, sizeof(double) * h->last_n * (size_t)h->last_nrec, hipMemcpyDeviceToHost));
    return PHC_OK;
}

// current != 0: level i2 (the field after the last step), else i1; node order.
phc_status phc_sem64_get_field(phc_sem64 *h, int current, double *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "NULL argument");
    PHC_HIP(hipStreamSynchronize(h->stream));
    const FdLayout &L = h->L;
    std::vector<double> buf((size_t)L.total);
    PHC_HIP(hipMemcpy(buf.data(), h->lev[current ? h->i2 : h->i1].ptr, sizeof(double) * L.total, hipMemcpyDeviceToHost));
    for (int iy = 0; iy < h->ny; ++iy)
        for (int iz = 0; iz < h->nz; ++iz)
            std::memcpy(host + ((int64_t)iy * h->nz + iz) * h->nx, &buf[(size_t)fd_off(L, iy, iz, 0)],
                        sizeof(double) * h->nx);
    return PHC_OK;
}

} // extern "C"
