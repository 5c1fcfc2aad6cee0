// fd64.hip -- FP64 FD time loop (phc_fd64_* C ABI): the reference's FD step
// (FDTDKernel::inner3D + applySponge + addRHS, src/fd/utilsFDTD/
// FDTDkernels.hpp:8-34, 108-128; time loop FDTDmain.cpp:55-81) with every
// field, coefficient and sponge factor in double.  It is the accuracy yardstick
// of the FP32 paths (e.g. the wide-domain run the learned boundary is judged
// against), not a throughput path: one thread per (k, j) column marching a
// chunk of i with the i-neighbours in a register queue, y/z neighbours from
// cache.  Each point's update is evaluated in the reference's order with no
// contraction; the sponge's damping of the current level is applied when that
// level is next read (the same products).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fd_kernels.hpp"
#include "phc_common.hpp"

namespace phc {

constexpr int F64_R = 4;
constexpr int F64_CHUNK = 16;

struct Fd64Args {
    const double *pn;
    double *pnm1;
    const double *vp;
    FdLayout L;
    double cx[F64_R + 1], cy[F64_R + 1], cz[F64_R + 1];
    double coef0;
    int sw;
    const double *fx, *fy, *fz;
    int damp_prev;
};

__device__ inline double f64_sponge(const Fd64Args &a, int i, int j, int k)
{
    if (a.sw == 0) return 1.0;
    const bool inI = i >= a.sw && i < a.L.nx - a.sw, inJ = j >= a.sw && j < a.L.ny - a.sw,
               inK = k >= a.sw && k < a.L.nz - a.sw;
    double s = 1.0;
    if (inJ && inK && !inI) s = a.fx[i];
    if (inK && inI && !inJ) s = a.fy[j];
    if (inJ && inI && !inK) s = a.fz[k];
    return s;
}

__global__ void __launch_bounds__(256) fd64_step_kernel(Fd64Args a)
{
#pragma clang fp contract(off)
    constexpr int R = F64_R;
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int i0 = blockIdx.z * F64_CHUNK;
    const FdLayout &L = a.L;
    if (k >= L.nz || j >= L.ny) return;
    const int i1 = min(i0 + F64_CHUNK, L.nx);
    const int64_t pj = L.pitch_j, pk = L.pitch_k;
    double q[2 * R + 1];
#pragma unroll
    for (int l = 0; l < 2 * R; ++l) q[l + 1] = a.pn[fd_off(L, i0 - R + l, j, k)];
    for (int i = i0; i < i1; ++i) {
#pragma unroll
        for (int l = 0; l < 2 * R; ++l) q[l] = q[l + 1];
        q[2 * R] = a.pn[fd_off(L, i + R, j, k)];
        const int64_t c = fd_off(L, i, j, k);
        const double pc = q[R];
        double lx = 0.0, ly = 0.0, lz = 0.0;
#pragma unroll
        for (int l = 1; l <= R; ++l) lx += a.cx[l] * (q[R + l] + q[R - l]);
#pragma unroll
        for (int l = 1; l <= R; ++l) ly += a.cy[l] * (a.pn[c + l * pk] + a.pn[c - l * pk]);
#pragma unroll
        for (int l = 1; l <= R; ++l) lz += a.cz[l] * (a.pn[c + l] + a.pn[c - l]);
        const double s = f64_sponge(a, i, j, k);
        const double pm = a.damp_prev ? a.pnm1[c] * s : a.pnm1[c];
        const double v = a.vp[c];
        a.pnm1[c] = (2.0 * pc - pm + v * (a.coef0 * pc + lx + ly + lz)) * s;
    }
}

__global__ void fd64_add_rhs_kernel(double *p, const double *vp, int64_t off, const double *rhs, int it)
{
#pragma clang fp contract(off)
    if (threadIdx.x == 0 && blockIdx.x == 0) p[off] += vp[off] * rhs[it];
}

__global__ void fd64_sample_kernel(const double *p, const int64_t *offs, int nrec, double *out, int stride,
                                   int step, int64_t src_off, double *src_trace)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) src_trace[step] = p[src_off];
    if (r < nrec) out[(int64_t)r * stride + step] = p[offs[r]];
}

} // namespace phc

using namespace phc;

struct phc_fd64 {
    FdLayout L;
    int xs, ys, zs, sw, nsamples;
    Fd64Args args;
    DevBuf lev[2], vp, fx, fy, fz, rhs, trace, rec, recoff;
    std::vector<double> hfx, hfy, hfz;
    int cur = 1, pending = 0;
    int nrec = 0, last_n = 0, last_nrec = 0;
    hipStream_t stream = nullptr;
};

static void f64_axis_damping(int n, int W, float alpha, std::vector<double> &f)
{
    f.assign(n, 1.0);
    for (int i = 0; i < n; ++i) {
        int d = -1;
        if (i >= n - W) d = W - (n - i);
        else if (i < W) d = W - i;
        if (d >= 0) f[i] = std::exp((double)alpha * std::pow((double)d, 2.0));
    }
}

template <class T>
static phc_status f64_upload(DevBuf &b, const std::vector<T> &v)
{
    PHC_TRY(b.alloc(sizeof(T) * std::max<size_t>(v.size(), 1)));
    if (!v.empty()) PHC_HIP(hipMemcpy(b.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return PHC_OK;
}

extern "C" {

phc_status phc_fd64_create(const phc_fd_params *p, phc_fd64 **out)
{
    if (!out) return fail(PHC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!p) return fail(PHC_ERR_INVALID, "params is NULL");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(PHC_ERR_INVALID, "empty grid");
    if (p->lx != F64_R || p->ly != F64_R || p->lz != F64_R)
        return fail(PHC_ERR_INVALID, "the FP64 path has half width %d only", F64_R);
    if (p->xs < 0 || p->xs >= p->nx || p->ys < 0 || p->ys >= p->ny || p->zs < 0 || p->zs >= p->nz)
        return fail(PHC_ERR_INVALID, "source outside the grid");
    if (!p->coefx || !p->coefy || !p->coefz || !p->vp || (p->nsamples > 0 && !p->rhs) || p->nsamples < 0)
        return fail(PHC_ERR_INVALID, "coefficient, vp or rhs array is NULL");
    if (p->sponge_width < 0 || p->sponge_width > 4096) return fail(PHC_ERR_INVALID, "bad sponge width");
    phc_fd64 *h = new phc_fd64();
    auto bail = [&](phc_status s) { phc_fd64_destroy(h); return s; };
    h->L = make_fd_layout(p->nx, p->ny, p->nz, F64_R);
    h->xs = p->xs; h->ys = p->ys; h->zs = p->zs;
    h->sw = p->sponge_width;
    h->nsamples = p->nsamples;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "hipStreamCreate failed"));
    const FdLayout &L = h->L;
    std::vector<double> buf((size_t)L.total, (double)p->init_value);
    phc_status st;
    if ((st = f64_upload(h->lev[0], buf)) || (st = f64_upload(h->lev[1], buf))) return bail(st);
    std::fill(buf.begin(), buf.end(), 0.0);
    for (int i = 0; i < L.nx; ++i)
        for (int j = 0; j < L.ny; ++j)
            for (int k = 0; k < L.nz; ++k)
                buf[(size_t)fd_off(L, i, j, k)] = (double)p->vp[((int64_t)i * L.ny + j) * L.nz + k];
    if ((st = f64_upload(h->vp, buf))) return bail(st);
    f64_axis_damping(L.nx, h->sw, p->sponge_alpha, h->hfx);
    f64_axis_damping(L.ny, h->sw, p->sponge_alpha, h->hfy);
    f64_axis_damping(L.nz, h->sw, p->sponge_alpha, h->hfz);
    if ((st = f64_upload(h->fx, h->hfx)) || (st = f64_upload(h->fy, h->hfy)) || (st = f64_upload(h->fz, h->hfz)))
        return bail(st);
    std::vector<double> rhs(p->rhs, p->rhs + p->nsamples);
    if ((st = f64_upload(h->rhs, rhs))) return bail(st);
    Fd64Args &a = h->args;
    std::memset(&a, 0, sizeof(a));
    a.L = L;
    a.vp = h->vp.as<double>();
    for (int l = 0; l <= F64_R; ++l) {
        a.cx[l] = p->coefx[l];
        a.cy[l] = p->coefy[l];
        a.cz[l] = p->coefz[l];
    }
    a.coef0 = p->coef0;
    a.sw = h->sw;
    a.fx = h->fx.as<double>();
    a.fy = h->fy.as<double>();
    a.fz = h->fz.as<double>();
    *out = h;
    return PHC_OK;
}

phc_status phc_fd64_destroy(phc_fd64 *h)
{
    if (!h) return PHC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (DevBuf *b : {&h->lev[0], &h->lev[1], &h->vp, &h->fx, &h->fy, &h->fz, &h->rhs, &h->trace, &h->rec,
                      &h->recoff})
        b->release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PHC_OK;
}

phc_status phc_fd64_set_receivers(phc_fd64 *h, int nrec, const int *ijk)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nrec < 0 || (nrec > 0 && !ijk)) return fail(PHC_ERR_INVALID, "bad receiver list");
    std::vector<int64_t> offs(nrec);
    for (int r = 0; r < nrec; ++r) {
        const int i = ijk[3 * r], j = ijk[3 * r + 1], k = ijk[3 * r + 2];
        if (i < 0 || i >= h->L.nx || j < 0 || j >= h->L.ny || k < 0 || k >= h->L.nz)
            return fail(PHC_ERR_INVALID, "receiver %d outside the grid", r);
        offs[r] = fd_off(h->L, i, j, k);
    }
    PHC_HIP(hipStreamSynchronize(h->stream));
    h->nrec = nrec;
    return f64_upload(h->recoff, offs);
}

// Samples it0 .. it0+nsteps-1: addRHS(current) -> step -> swap, as FDTDmain.cpp:55-81.
phc_status phc_fd64_run(phc_fd64 *h, int it0, int nsteps)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nsteps < 0 || it0 < 0 || it0 + nsteps > h->nsamples)
        return fail(PHC_ERR_INVALID, "steps %d..%d outside the %d samples", it0, it0 + nsteps - 1, h->nsamples);
    if (nsteps == 0) return PHC_OK;
    PHC_TRY(h->trace.alloc(sizeof(double) * nsteps));
    if (h->nrec > 0) PHC_TRY(h->rec.alloc(sizeof(double) * (size_t)nsteps * h->nrec));
    const FdLayout &L = h->L;
    const int64_t src = fd_off(L, h->xs, h->ys, h->zs);
    dim3 grid((L.nz + 63) / 64, (L.ny + 3) / 4, (L.nx + F64_CHUNK - 1) / F64_CHUNK);
    for (int s = 0; s < nsteps; ++s) {
        const int cb = h->cur, ca = 1 - h->cur;
        hipLaunchKernelGGL(fd64_add_rhs_kernel, dim3(1), dim3(64), 0, h->stream, h->lev[cb].as<double>(),
                           h->vp.as<double>(), src, h->rhs.as<double>(), it0 + s);
        Fd64Args a = h->args;
        a.pn = h->lev[cb].as<double>();
        a.pnm1 = h->lev[ca].as<double>();
        a.damp_prev = h->pending;
        hipLaunchKernelGGL(fd64_step_kernel, grid, dim3(64, 4), 0, h->stream, a);
        hipLaunchKernelGGL(fd64_sample_kernel, dim3(std::max(1, (h->nrec + 255) / 256)), dim3(256), 0, h->stream,
                           h->lev[ca].as<double>(), h->recoff.as<int64_t>(), h->nrec, h->rec.as<double>(), nsteps,
                           s, src, h->trace.as<double>());
        PHC_HIP(hipGetLastError());
        h->cur = ca;
        h->pending = h->sw > 0 ? 1 : 0;
    }
    h->last_n = nsteps;
    h->last_nrec = h->nrec;
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

phc_status phc_fd64_get_traces(phc_fd64 *h, double *src, double *rec)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    if (src && h->last_n) PHC_HIP(hipMemcpy(src, h->trace.ptr, sizeof(double) * h->last_n, hipMemcpyDeviceToHost));
    if (rec && h->last_n && h->last_nrec)
        PHC_HIP(hipMemcpy(rec, h->rec.ptr, sizeof(double) * h->last_n * (size_t)h->last_nrec, hipMemcpyDeviceToHost));
    return PHC_OK;
}

// current != 0: the current level, else the previous one (with its pending
// sponge product applied), extended layout as phc_fd_get_field.
phc_status phc_fd64_get_field(phc_fd64 *h, int current, double *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "NULL argument");
    PHC_HIP(hipStreamSynchronize(h->stream));
    const FdLayout &L = h->L;
    const int level = current ? h->cur : 1 - h->cur;
    std::vector<double> buf((size_t)L.total);
    PHC_HIP(hipMemcpy(buf.data(), h->lev[level].ptr, sizeof(double) * L.total, hipMemcpyDeviceToHost));
    const int R = F64_R, NYe = L.ny + 2 * R, NZe = L.nz + 2 * R;
    const bool damp = !current && h->pending;
    for (int i = -R; i < L.nx + R; ++i)
        for (int j = -R; j < L.ny + R; ++j) {
            double *dst = host + ((int64_t)(i + R) * NYe + (j + R)) * NZe;
            std::memcpy(dst, &buf[(size_t)fd_off(L, i, j, -R)], sizeof(double) * NZe);
            if (!damp || i < 0 || i >= L.nx || j < 0 || j >= L.ny) continue;
            for (int k = 0; k < L.nz; ++k) {
                const int W = h->sw;
                const bool inI = i >= W && i < L.nx - W, inJ = j >= W && j < L.ny - W, inK = k >= W && k < L.nz - W;
                double s = 1.0;
                if (inJ && inK && !inI) s = h->hfx[i];
                if (inK && inI && !inJ) s = h->hfy[j];
                if (inJ && inI && !inK) s = h->hfz[k];
                dst[k + R] *= s;
            }
        }
    return PHC_OK;
}

} // extern "C"
