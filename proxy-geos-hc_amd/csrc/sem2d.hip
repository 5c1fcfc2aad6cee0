// sem2d.hip -- the SEM step of the reference's SEM2D build and the
// phc_sem2d_* C ABI.
//
// Reference: SEMsolver::computeOneStep with SEM2D defined
// (src/sem/solver/SEMsolver.cpp:21-125): ex x ez quadrilaterals of order p on
// the x-z plane (SEMmesh with ey = 0, SEMmesh.cpp:3-26), leapfrog on the
// interior nodes and, on the four edges, the first-order paraxial condition
//   M pnew = 2 M p - (M - dt S/2) pold - dt^2 y,  divided by (M + dt S/2),
// with S the boundary mass of 1/c (computeDs, SEMQkGLIntegrals.hpp:568-599;
// SEMsolver.cpp:91-123).
//
// On the axis-aligned mesh the assembled stiffness is Ax (x) Wz + Wx (x) Az
// (1D assembled operators A, lumped 1D masses W), so with Q = dt^2 Wx Wz / M
// and sigma = dt S / (2 M) every node's update is
//   pnew = (2 p - (1 - sigma) pold - Q (Lx p + Lz p)) / (1 + sigma),
// Lx p = (Ax p)/Wx a radius-p stencil whose row depends only on the node's
// position class along the axis: first node, last node, or index mod p.
// Rows sum to zero, and the stencil is evaluated on differences from the
// centre node: in FP32 that is about 17x closer to an FP64 run than the plain
// sum on the reference's published 2D case (6.6e-6 against 1.1e-4 relative
// L2 after 900 steps; the reference's own FP32 element loop: 7.6e-5).
// sigma is non-zero on the edges only and is read from four 1D arrays.
//
// Kernel: one wavefront owns 64 columns and marches along z over a chunk of
// rows; the z-neighbours sit in a register ring, the current row (plus p
// halo columns each side) is staged in a per-wave LDS row for the
// x-neighbours.  No workgroup barrier: each wave only reads the LDS row it
// wrote; wavefront-scope fences order those accesses.  HBM traffic per node
// update: p, pold, Q read, pnew written (16 B).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "march.hpp"
#include "phc_common.hpp"
#include "sem_setup.hpp"

namespace phc {

constexpr int S2_MAX_P = 5;
constexpr int S2_W = 2 * S2_MAX_P + 1;
constexpr int S2_WAVES = 4;  // waves per block
constexpr int S2_DEPTH = 2;  // rows prefetched ahead of the one computed

// 2D layout: node (ix, iz), ix in [-p, tiles_x*64 + p), iz in [-p, nz + p),
// at origin + iz*pitch + ix; interior ix = 0 starts 32 floats into a
// 128-byte-aligned row.
struct Sem2Layout {
    int nx, nz, p, tiles_x;
    int64_t pitch, origin, total;
};

inline Sem2Layout make_sem2_layout(int nx, int nz, int p)
{
    Sem2Layout L;
    L.nx = nx; L.nz = nz; L.p = p;
    L.tiles_x = (nx + 63) / 64;
    L.pitch = round_up(32 + (int64_t)L.tiles_x * 64 + p, 32);
    L.origin = (int64_t)p * L.pitch + 32;
    L.total = (int64_t)(nz + 2 * p) * L.pitch;
    return L;
}

struct Sem2dArgs {
    const float *pn;    // level i2 (current)
    float *pnm1;        // level i1: read, overwritten by the new level
    const float *qf;    // dt^2 Wx Wz / M per node
    const float *sig_z0, *sig_z1;  // sigma on rows iz = 0 and iz = nz-1, by ix
    const float *sig_x0, *sig_x1;  // sigma on columns ix = 0 and ix = nx-1, by iz
    Sem2Layout L;
    // stencil rows by position class: 0..p-1 = index mod p (interior),
    // p = first node, p+1 = last node; offsets -p..p
    float cx[S2_MAX_P + 2][S2_W];
    float cz[S2_MAX_P + 2][S2_W];
    int64_t src_off;    // layout offset of the source node
    int inj;
    float inj_coef;     // dt^2 c^2 of the source element (SEMsolver.cpp:41-42)
    float rhs_next;
    float *trace;
    int chunk, nchunks;
};

template <int P, typename IT>
__global__ void __launch_bounds__(64 * S2_WAVES) sem2d_step_kernel(Sem2dArgs a)
{
    constexpr int W = 2 * P + 1;
    constexpr int RW = 64 + 2 * P;
    constexpr IT S = sizeof(IT) == 4 ? 4 : 1;
    __shared__ float srow[S2_WAVES][RW];

    const Sem2Layout &L = a.L;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wave = blockIdx.x * S2_WAVES + wv;
    if (wave >= L.tiles_x * a.nchunks) return;  // whole wave; no block barrier below
    const int tx = wave % L.tiles_x, ch = wave / L.tiles_x;
    const int x0 = tx * 64, x = x0 + lane;
    const int z0 = ch * a.chunk, z1 = min(z0 + a.chunk, L.nz);
    const bool xin = x < L.nx;
    const IT pitch = (IT)L.pitch * S;
    float *row = srow[wv];

    // per-lane x stencil row (the column does not change along the march)
    float cxr[W];
    {
        const int cls = (x == 0) ? P : (x == L.nx - 1) ? P + 1 : (xin ? x % P : 0);
#pragma unroll
        for (int d = 0; d < W; ++d) cxr[d] = a.cx[cls][d];
    }
    // halo columns: lanes 0..p-1 -> x0-p+lane, lanes p..2p-1 -> x0+64+lane-p
    const bool hl = lane < 2 * P;
    const int hx = (lane < P) ? (x0 - P + lane) : (x0 + 64 + lane - P);
    const int hcol = (lane < P) ? lane : (64 + lane);
    const IT col = (IT)(L.origin + x) * S;
    const IT hcolo = (IT)(L.origin + (hl ? hx : x)) * S;

    float rz[W];  // rz[d] = p at row z - P + d
#pragma unroll
    for (int d = 0; d < 2 * P; ++d) rz[d + 1] = ldo(a.pn, col + (IT)(z0 - P + d) * pitch);
    // rows z+1 .. z+D are in flight while row z is computed: slot s of the
    // buffers holds row zb + s (the ring entry P rows further down, the halo
    // columns, pold and Q); rows are clamped inside the array
    float nb[S2_DEPTH], hb[S2_DEPTH], pb[S2_DEPTH], qb[S2_DEPTH];
#pragma unroll
    for (int s = 0; s < S2_DEPTH; ++s) {
        const IT zr = (IT)min(z0 + s, L.nz - 1) * pitch;
        nb[s] = ldo(a.pn, col + (IT)min(z0 + s + P, L.nz - 1 + P) * pitch);
        hb[s] = ldo(a.pn, hcolo + zr);
        pb[s] = ldo_nt(a.pnm1, col + zr);
        qb[s] = ldo_nt(a.qf, col + zr);
    }

    for (int zb = z0; zb < z1; zb += S2_DEPTH) {
#pragma unroll
        for (int s = 0; s < S2_DEPTH; ++s) {
            const int z = zb + s;
            if (z >= z1) break;
#pragma unroll
            for (int d = 0; d < 2 * P; ++d) rz[d] = rz[d + 1];
            rz[2 * P] = nb[s];
            const float hcur = hb[s], po = pb[s], q = qb[s];
            {
                const IT zn = (IT)min(z + S2_DEPTH, L.nz - 1) * pitch;
                nb[s] = ldo(a.pn, col + (IT)min(z + S2_DEPTH + P, L.nz - 1 + P) * pitch);
                hb[s] = ldo(a.pn, hcolo + zn);
                pb[s] = ldo_nt(a.pnm1, col + zn);
                qb[s] = ldo_nt(a.qf, col + zn);
            }
            const IT zo = (IT)z * pitch;

            row[P + lane] = rz[P];
            if (hl) row[hcol] = hcur;
            wave_lds_sync();
            const int cz = (z == 0) ? P : (z == L.nz - 1) ? P + 1 : z % P;
            // every stencil row sums to zero (the operator annihilates
            // constants), so the sum is taken over differences from the centre
            // value: the large common part cancels exactly and FP32 keeps the
            // small differences a smooth field produces
            const float pc = rz[P];
            float lap = 0.f;
#pragma unroll
            for (int d = 0; d < W; ++d) {
                if (d == P) continue;
                lap = fmaf(a.cz[cz][d], rz[d] - pc, lap);
                lap = fmaf(cxr[d], row[lane + d] - pc, lap);
            }
            wave_lds_sync();
            if (!xin) continue;
            float val;
            const bool edge = z == 0 || z == L.nz - 1 || x == 0 || x == L.nx - 1;
            if (!edge) {
                val = fmaf(-q, lap, 2.f * pc - po);
            } else {
                const float sg = (z == 0) ? a.sig_z0[x] : (z == L.nz - 1) ? a.sig_z1[x]
                               : (x == 0) ? a.sig_x0[z] : a.sig_x1[z];
                val = fmaf(-q, lap, fmaf(1.f - sg, -po, 2.f * pc)) / (1.f + sg);
            }
            const IT o = col + zo;
            if ((int64_t)(o / S) == a.src_off) {
                if (a.trace) *a.trace = val;
                if (a.inj) val = add_nc(val, mul_nc(a.inj_coef, a.rhs_next));
            }
            sto_nt(a.pnm1, o, val);
        }
    }
}

__global__ void sem2d_inject_kernel(float *p, int64_t off, float coef, const float *rhs, int it)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[off] = add_nc(p[off], mul_nc(coef, rhs[it]));
}

// One source term handed over by the caller: p[off] += coef * value.
__global__ void sem2d_add_value_kernel(float *p, int64_t off, float coef, float value)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[off] = add_nc(p[off], mul_nc(coef, value));
}

// Receiver samples after a step.  The source node already holds the next
// sample's injection, so a receiver there takes the source trace (the value
// before injection) instead.
__global__ void sem2d_sample_kernel(const float *p, const int64_t *offs, int nrec, float *out, int stride,
                                    int step, int64_t src_off, const float *src_trace)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int64_t o = offs[r];
    out[(int64_t)r * stride + step] = (o == src_off) ? src_trace[step] : p[o];
}

using Sem2dFn = void (*)(Sem2dArgs);

template <int P>
static Sem2dFn s2_idx(bool idx32)
{
    return idx32 ? sem2d_step_kernel<P, uint32_t> : sem2d_step_kernel<P, int64_t>;
}

static Sem2dFn sem2d_pick(int p, bool idx32)
{
    switch (p) {
    case 1: return s2_idx<1>(idx32);
    case 2: return s2_idx<2>(idx32);
    case 3: return s2_idx<3>(idx32);
    case 4: return s2_idx<4>(idx32);
    case 5: return s2_idx<5>(idx32);
    default: return nullptr;
    }
}

} // namespace phc

// ====================================================================== API
using namespace phc;

struct phc_sem2d {
    int ex, ez, p, nx, nz;
    int64_t nn;
    float dt;
    int nsamples;
    std::vector<float> hrhs, model;
    int src_element, src_node;
    float inj_coef;
    hipStream_t stream = nullptr;
    Sem2Layout L;
    DevBuf lev[2], qf, sig, rhs;   // sig: z0[nx], z1[nx], x0[nz], x1[nz]
    int index64 = 0;               // 64-bit offsets even for arrays < 4 GiB
    float cx[S2_MAX_P + 2][S2_W], cz[S2_MAX_P + 2][S2_W];
    std::vector<float> hmass, hsh;  // FP32 lumped mass and boundary term per node
    int chunk = 0;
    DevBuf trace, rec, recoff;
    int nrec = 0, last_n = 0, last_nrec = 0;
    int timing = 0;
    double last_ms = 0.0;
    hipGraphExec_t gexec = nullptr;
    int g_it0 = -1, g_n = -1, g_i1 = -1, g_nrec = -1;
};

static int64_t s2_off(const phc_sem2d *h, int64_t node)
{
    const int ix = (int)(node % h->nx), iz = (int)(node / h->nx);
    return h->L.origin + (int64_t)iz * h->L.pitch + ix;
}

static void s2_drop_graph(phc_sem2d *h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    h->g_it0 = h->g_n = h->g_i1 = h->g_nrec = -1;
}

static phc_status s2_put(phc_sem2d *h, DevBuf &dst, const float *host)
{
    std::vector<float> buf((size_t)h->L.total, 0.f);
    for (int iz = 0; iz < h->nz; ++iz)
        std::memcpy(&buf[(size_t)s2_off(h, (int64_t)iz * h->nx)], host + (int64_t)iz * h->nx,
                    sizeof(float) * h->nx);
    PHC_HIP(hipMemcpy(dst.ptr, buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice));
    return PHC_OK;
}

static phc_status s2_get(phc_sem2d *h, const DevBuf &src, float *host)
{
    PHC_HIP(hipStreamSynchronize(h->stream));
    std::vector<float> buf((size_t)h->L.total);
    PHC_HIP(hipMemcpy(buf.data(), src.ptr, sizeof(float) * buf.size(), hipMemcpyDeviceToHost));
    for (int iz = 0; iz < h->nz; ++iz)
        std::memcpy(host + (int64_t)iz * h->nx, &buf[(size_t)s2_off(h, (int64_t)iz * h->nx)],
                    sizeof(float) * h->nx);
    return PHC_OK;
}

static int s2_default_chunk(const phc_sem2d *h)
{
    // enough waves to fill the chip (about 4 per SIMD), chunks of at least 8 rows
    const int64_t want = 4096;
    int c = (int)std::max<int64_t>(8, (int64_t)h->L.tiles_x * h->nz / want);
    return std::min(c, h->nz);
}

static phc_status s2_launch(phc_sem2d *h, int i1, int i2, int inj, float rhs_next, float *trace,
                            hipEvent_t e0, hipEvent_t e1)
{
    Sem2dArgs a;
    std::memset(&a, 0, sizeof(a));
    a.pn = h->lev[i2].as<float>();
    a.pnm1 = h->lev[i1].as<float>();
    a.qf = h->qf.as<float>();
    float *sg = h->sig.as<float>();
    a.sig_z0 = sg;
    a.sig_z1 = sg + h->nx;
    a.sig_x0 = sg + 2 * h->nx;
    a.sig_x1 = sg + 2 * h->nx + h->nz;
    a.L = h->L;
    std::memcpy(a.cx, h->cx, sizeof(a.cx));
    std::memcpy(a.cz, h->cz, sizeof(a.cz));
    a.src_off = s2_off(h, h->src_node);
    a.inj = inj;
    a.inj_coef = h->inj_coef;
    a.rhs_next = rhs_next;
    a.trace = trace;
    a.chunk = h->chunk > 0 ? h->chunk : s2_default_chunk(h);
    a.nchunks = (h->nz + a.chunk - 1) / a.chunk;
    const int waves = h->L.tiles_x * a.nchunks;
    Sem2dFn fn = sem2d_pick(h->p, !h->index64 && h->L.total * 4 < ((int64_t)1 << 32));
    if (!fn) return fail(PHC_ERR_INVALID, "no 2D SEM kernel for p=%d", h->p);
    if (e0) PHC_HIP(hipEventRecord(e0, h->stream));
    hipLaunchKernelGGL(fn, dim3((waves + S2_WAVES - 1) / S2_WAVES), dim3(64 * S2_WAVES), 0, h->stream, a);
    PHC_HIP(hipGetLastError());
    if (e1) PHC_HIP(hipEventRecord(e1, h->stream));
    return PHC_OK;
}

static phc_status s2_inject(phc_sem2d *h, int level, int it)
{
    hipLaunchKernelGGL(sem2d_inject_kernel, dim3(1), dim3(64), 0, h->stream, h->lev[level].as<float>(),
                       s2_off(h, h->src_node), h->inj_coef, h->rhs.as<float>(), it);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

extern "C" {

phc_status phc_sem2d_destroy(phc_sem2d *h)
{
    if (!h) return PHC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    s2_drop_graph(h);
    for (DevBuf *b : {&h->lev[0], &h->lev[1], &h->qf, &h->sig, &h->rhs, &h->trace, &h->rec, &h->recoff})
        b->release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PHC_OK;
}

phc_status phc_sem2d_create(const phc_sem_params *p, phc_sem2d **out)
{
    if (!out) return fail(PHC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!p) return fail(PHC_ERR_INVALID, "params is NULL");
    if (p->ey != 0) return fail(PHC_ERR_INVALID, "a 2D mesh has ey = 0 (got %d)", p->ey);
    if (p->ex < 2 || p->ez < 2) return fail(PHC_ERR_INVALID, "2D mesh %d x %d: need at least 2 x 2 elements", p->ex, p->ez);
    if (p->order < 1 || p->order > S2_MAX_P)
        return fail(PHC_ERR_INVALID, "order %d outside 1..%d", p->order, S2_MAX_P);
    if (!(p->lx > 0) || !(p->lz > 0)) return fail(PHC_ERR_INVALID, "bad domain size");
    if (!p->model) return fail(PHC_ERR_INVALID, "model is NULL");
    if (p->nsamples < 0 || (p->nsamples > 0 && !p->rhs)) return fail(PHC_ERR_INVALID, "bad rhs");
    if (p->kernel != PHC_SEM_SEPARABLE)
        return fail(PHC_ERR_INVALID, "the 2D step has the separable kernel only");
    if (!(p->dt > 0)) return fail(PHC_ERR_INVALID, "dt must be positive");
    const int64_t nE = (int64_t)p->ex * p->ez;
    if (p->src_element < 0 || p->src_element >= nE)
        return fail(PHC_ERR_INVALID, "source element %d outside 0..%lld", p->src_element, (long long)nE - 1);
    for (int64_t e = 0; e < nE; ++e)
        if (!(p->model[e] > 0.f) || !std::isfinite(p->model[e]))
            return fail(PHC_ERR_INVALID, "wave speed of element %lld is not positive", (long long)e);
    // element sizes as the reference's coordinate routine sees them: the
    // float quotient (SEMmesh.cpp:20-22) through an int parameter (SEMmesh.hpp:61)
    const int hxi = (int)(float)((double)p->lx / (1. * p->ex));
    const int hzi = (int)(float)((double)p->lz / (1. * p->ez));
    if (hxi < 1 || hzi < 1) return fail(PHC_ERR_INVALID, "element size below 1 (the reference truncates it to an int)");
    const int64_t nx = (int64_t)p->ex * p->order + 1, nz = (int64_t)p->ez * p->order + 1;
    if (nx * nz > ((int64_t)1 << 31) - 1) return fail(PHC_ERR_INVALID, "mesh has more than 2^31 nodes");

    phc_sem2d *h = new phc_sem2d();
    auto bail = [&](phc_status s) { phc_sem2d_destroy(h); return s; };
    h->ex = p->ex; h->ez = p->ez; h->p = p->order;
    h->nx = (int)nx; h->nz = (int)nz;
    h->nn = nx * nz;
    h->dt = p->dt;
    h->nsamples = p->nsamples;
    h->hrhs.assign(p->rhs, p->rhs + p->nsamples);
    h->model.assign(p->model, p->model + (size_t)p->ex * p->ez);
    h->src_element = p->src_element;
    {
        const int e = p->src_element, exi = e % h->ex, ezi = e / h->ex;
        h->src_node = exi * h->p + ezi * h->p * h->nx;  // first node of the element
        const float c = p->model[e];
        h->inj_coef = ((h->dt * h->dt) * c) * c;         // SEMsolver.cpp:41-42
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "hipStreamCreate failed"));

    // ---- FP64 setup: 1D operators, lumped mass, boundary term
    const int P = h->p, qn = P + 1, W = 2 * P + 1;
    const GllBasis g = make_gll(P, p->weights_ref != 0);
    const SemAxis ax = make_axis(g, h->ex, hxi), az = make_axis(g, h->ez, hzi);
    std::vector<double> M((size_t)h->nn, 0.0), Sh((size_t)h->nn, 0.0);
    for (int ezi = 0; ezi < h->ez; ++ezi)
        for (int exi = 0; exi < h->ex; ++exi) {
            const double c = p->model[exi + (int64_t)ezi * h->ex];
            const double ic2 = 1.0 / (c * c);
            for (int b = 0; b < qn; ++b)
                for (int a2 = 0; a2 < qn; ++a2) {
                    const int64_t node = (exi * P + a2) + (int64_t)(ezi * P + b) * h->nx;
                    M[node] += (g.w[a2] * ax.half) * (g.w[b] * az.half) * ic2;
                }
            // paraxial boundary mass: w_a (h/2) / c on each boundary face
            // (SEMsolver.cpp:95-110; ds = h/2 on the box)
            for (int a2 = 0; a2 < qn; ++a2) {
                if (ezi == 0) Sh[exi * P + a2] += g.w[a2] * ax.half / c;
                if (ezi == h->ez - 1) Sh[exi * P + a2 + (int64_t)(h->nz - 1) * h->nx] += g.w[a2] * ax.half / c;
                if (exi == 0) Sh[(int64_t)(ezi * P + a2) * h->nx] += g.w[a2] * az.half / c;
                if (exi == h->ex - 1) Sh[(int64_t)(ezi * P + a2) * h->nx + h->nx - 1] += g.w[a2] * az.half / c;
            }
        }
    h->hmass.resize(h->nn);
    h->hsh.resize(h->nn);
    for (int64_t n = 0; n < h->nn; ++n) {
        h->hmass[n] = (float)M[n];
        h->hsh[n] = (float)Sh[n];
    }
    const double dt2 = (double)h->dt * (double)h->dt;
    h->L = make_sem2_layout(h->nx, h->nz, P);
    phc_status st;
    const size_t bytes = sizeof(float) * (size_t)h->L.total;
    if ((st = h->lev[0].alloc(bytes)) || (st = h->lev[1].alloc(bytes)) || (st = h->qf.alloc(bytes)) ||
        (st = h->sig.alloc(sizeof(float) * 2 * (size_t)(h->nx + h->nz))) ||
        (st = h->rhs.alloc(sizeof(float) * std::max(1, h->nsamples))))
        return bail(st);
    if (hipMemset(h->lev[0].ptr, 0, bytes) || hipMemset(h->lev[1].ptr, 0, bytes))
        return bail(fail(PHC_ERR_HIP, "level clear failed"));
    if (h->nsamples > 0 && hipMemcpy(h->rhs.ptr, p->rhs, sizeof(float) * h->nsamples, hipMemcpyHostToDevice))
        return bail(fail(PHC_ERR_HIP, "rhs upload failed"));
    {
        std::vector<float> Qn((size_t)h->nn);
        for (int iz = 0; iz < h->nz; ++iz)
            for (int ix = 0; ix < h->nx; ++ix) {
                const int64_t n = ix + (int64_t)iz * h->nx;
                Qn[n] = (float)(dt2 * ax.Wg[ix] * az.Wg[iz] / M[n]);
            }
        if ((st = s2_put(h, h->qf, Qn.data()))) return bail(st);
        auto sigma = [&](int64_t n) { return (float)((double)h->dt * Sh[n] / (2.0 * M[n])); };
        std::vector<float> sg(2 * (size_t)(h->nx + h->nz));
        for (int ix = 0; ix < h->nx; ++ix) {
            sg[ix] = sigma(ix);
            sg[h->nx + ix] = sigma(ix + (int64_t)(h->nz - 1) * h->nx);
        }
        for (int iz = 0; iz < h->nz; ++iz) {
            sg[2 * h->nx + iz] = sigma((int64_t)iz * h->nx);
            sg[2 * h->nx + h->nz + iz] = sigma((int64_t)iz * h->nx + h->nx - 1);
        }
        if (hipMemcpy(h->sig.ptr, sg.data(), sizeof(float) * sg.size(), hipMemcpyHostToDevice))
            return bail(fail(PHC_ERR_HIP, "boundary upload failed"));
    }
    // stencil rows by position class (uniform element size along each axis,
    // so interior nodes with equal index mod p share a row)
    std::memset(h->cx, 0, sizeof(h->cx));
    std::memset(h->cz, 0, sizeof(h->cz));
    auto rows = [&](const SemAxis &A, float (*c)[S2_W]) {
        for (int m = 0; m < P; ++m) {
            int i = (m == 0) ? P : m;  // an interior node with residue m (ex >= 2)
            for (int d = 0; d < W; ++d) c[m][d] = (float)A.A[(size_t)i * W + d];
        }
        for (int d = 0; d < W; ++d) {
            c[P][d] = (float)A.A[(size_t)0 * W + d];
            c[P + 1][d] = (float)A.A[(size_t)(A.n - 1) * W + d];
        }
    };
    rows(ax, h->cx);
    rows(az, h->cz);
    h->chunk = p->chunk_i > 0 ? std::min(p->chunk_i, h->nz) : 0;
    h->index64 = p->index64 ? 1 : 0;
    *out = h;
    return PHC_OK;
}

phc_status phc_sem2d_sizes(phc_sem2d *h, int *num_nodes, int *nx, int *nz)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (num_nodes) *num_nodes = (int)h->nn;
    if (nx) *nx = h->nx;
    if (nz) *nz = h->nz;
    return PHC_OK;
}

phc_status phc_sem2d_compute_one_step(phc_sem2d *h, int time_sample, int i1, int i2)
{
    RoctxRange rr_("phc_sem2d_compute_one_step");
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (i1 < 0 || i1 > 1 || i2 < 0 || i2 > 1 || i1 == i2)
        return fail(PHC_ERR_INVALID, "levels i1=%d i2=%d must be 0 and 1", i1, i2);
    if (time_sample < 0 || time_sample >= h->nsamples)
        return fail(PHC_ERR_INVALID, "time sample %d outside 0..%d", time_sample, h->nsamples - 1);
    PHC_TRY(s2_inject(h, i2, time_sample));
    return s2_launch(h, i1, i2, 0, 0.f, nullptr, nullptr, nullptr);
}

phc_status phc_sem2d_compute_one_step_rhs(phc_sem2d *h, int i1, int i2, int nrhs, const int *rhs_element,
                                          const float *rhs_value)
{
    RoctxRange rr_("phc_sem2d_compute_one_step_rhs");
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (i1 < 0 || i1 > 1 || i2 < 0 || i2 > 1 || i1 == i2)
        return fail(PHC_ERR_INVALID, "levels i1=%d i2=%d must be 0 and 1", i1, i2);
    if (nrhs < 0 || (nrhs > 0 && (!rhs_element || !rhs_value)))
        return fail(PHC_ERR_INVALID, "bad source list (%d sources)", nrhs);
    const int nE = h->ex * h->ez;
    for (int r = 0; r < nrhs; ++r) {
        if (rhs_element[r] < 0 || rhs_element[r] >= nE)
            return fail(PHC_ERR_INVALID, "source %d: element %d outside 0..%d", r, rhs_element[r], nE - 1);
        if (!std::isfinite(rhs_value[r])) return fail(PHC_ERR_INVALID, "source %d: value is not finite", r);
    }
    // SEMsolver.cpp:37-42: pnGlobal(node0(e), i2) += dt dt c c rhs, source by source
    for (int r = 0; r < nrhs; ++r) {
        const int e = rhs_element[r];
        const int64_t node = (int64_t)(e % h->ex) * h->p + (int64_t)(e / h->ex) * h->p * h->nx;
        const float c = h->model[e];
        hipLaunchKernelGGL(sem2d_add_value_kernel, dim3(1), dim3(64), 0, h->stream, h->lev[i2].as<float>(),
                           s2_off(h, node), ((h->dt * h->dt) * c) * c, rhs_value[r]);
        PHC_HIP(hipGetLastError());
    }
    return s2_launch(h, i1, i2, 0, 0.f, nullptr, nullptr, nullptr);
}

phc_status phc_sem2d_set_source(phc_sem2d *h, int src_element, int nsamples, const float *rhs)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (src_element < 0 || src_element >= h->ex * h->ez)
        return fail(PHC_ERR_INVALID, "source element %d outside 0..%d", src_element, h->ex * h->ez - 1);
    if (nsamples < 0 || (nsamples > 0 && !rhs)) return fail(PHC_ERR_INVALID, "bad rhs");
    PHC_HIP(hipStreamSynchronize(h->stream));
    s2_drop_graph(h);
    h->src_element = src_element;
    h->src_node = (src_element % h->ex) * h->p + (src_element / h->ex) * h->p * h->nx;
    const float c = h->model[src_element];
    h->inj_coef = ((h->dt * h->dt) * c) * c;
    h->nsamples = nsamples;
    h->hrhs.assign(rhs, rhs + nsamples);
    PHC_TRY(h->rhs.alloc(sizeof(float) * std::max(1, nsamples)));
    if (nsamples > 0) PHC_HIP(hipMemcpy(h->rhs.ptr, rhs, sizeof(float) * nsamples, hipMemcpyHostToDevice));
    return PHC_OK;
}

static phc_status s2_enqueue_run(phc_sem2d *h, int it0, int nsteps, int i1, int i2, std::vector<hipEvent_t> *ev)
{
    PHC_TRY(s2_inject(h, i2, it0));
    for (int s = 0; s < nsteps; ++s) {
        const int it = it0 + s;
        const bool more = s + 1 < nsteps;
        PHC_TRY(s2_launch(h, i1, i2, more ? 1 : 0, more ? h->hrhs[it + 1] : 0.f, h->trace.as<float>() + s,
                          ev ? (*ev)[2 * s] : nullptr, ev ? (*ev)[2 * s + 1] : nullptr));
        if (h->nrec > 0) {
            hipLaunchKernelGGL(sem2d_sample_kernel, dim3((h->nrec + 255) / 256), dim3(256), 0, h->stream,
                               h->lev[i1].as<float>(), h->recoff.as<int64_t>(), h->nrec, h->rec.as<float>(),
                               nsteps, s, s2_off(h, h->src_node), h->trace.as<float>());
            PHC_HIP(hipGetLastError());
        }
        std::swap(i1, i2);
    }
    return PHC_OK;
}

phc_status phc_sem2d_run(phc_sem2d *h, int it0, int nsteps, int *i1p, int *i2p)
{
    RoctxRange rr_("phc_sem2d_run");
    if (!h || !i1p || !i2p) return fail(PHC_ERR_INVALID, "NULL argument");
    const int i1 = *i1p, i2 = *i2p;
    if (i1 < 0 || i1 > 1 || i2 < 0 || i2 > 1 || i1 == i2)
        return fail(PHC_ERR_INVALID, "levels i1=%d i2=%d must be 0 and 1", i1, i2);
    if (nsteps < 0) return fail(PHC_ERR_INVALID, "negative step count");
    if (nsteps == 0) return PHC_OK;
    if (it0 < 0 || it0 + nsteps > h->nsamples)
        return fail(PHC_ERR_INVALID, "steps %d..%d outside the %d source samples", it0, it0 + nsteps - 1,
                    h->nsamples);
    if (h->trace.bytes < sizeof(float) * nsteps) {
        PHC_HIP(hipStreamSynchronize(h->stream));
        s2_drop_graph(h);
        PHC_TRY(h->trace.alloc(sizeof(float) * nsteps));
    }
    if (h->nrec > 0 && h->rec.bytes < sizeof(float) * (size_t)nsteps * h->nrec) {
        PHC_HIP(hipStreamSynchronize(h->stream));
        s2_drop_graph(h);
        PHC_TRY(h->rec.alloc(sizeof(float) * (size_t)nsteps * h->nrec));
    }
    if (h->timing) {
        std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
        for (auto &e : ev) PHC_HIP(hipEventCreate(&e));
        phc_status st = s2_enqueue_run(h, it0, nsteps, i1, i2, &ev);
        if (st == PHC_OK) {
            PHC_HIP(hipStreamSynchronize(h->stream));
            double tot = 0.0;
            for (int s = 0; s < nsteps; ++s) {
                float ms = 0.f;
                PHC_HIP(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
                tot += ms;
            }
            h->last_ms = tot / nsteps;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        PHC_TRY(st);
    } else {
        if (!(h->gexec && h->g_it0 == it0 && h->g_n == nsteps && h->g_i1 == i1 && h->g_nrec == h->nrec)) {
            s2_drop_graph(h);
            hipGraph_t graph = nullptr;
            PHC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            phc_status st = s2_enqueue_run(h, it0, nsteps, i1, i2, nullptr);
            hipError_t ce = hipStreamEndCapture(h->stream, &graph);
            if (st != PHC_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return st;
            }
            if (ce != hipSuccess) return fail(PHC_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                h->gexec = nullptr;
                return fail(PHC_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
            }
            h->g_it0 = it0; h->g_n = nsteps; h->g_i1 = i1; h->g_nrec = h->nrec;
        }
        PHC_HIP(hipGraphLaunch(h->gexec, h->stream));
    }
    h->last_n = nsteps;
    h->last_nrec = h->nrec;
    if (nsteps % 2) std::swap(*i1p, *i2p);
    return PHC_OK;
}

phc_status phc_sem2d_set_receivers(phc_sem2d *h, int nrec, const int *nodes)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nrec < 0 || (nrec > 0 && !nodes)) return fail(PHC_ERR_INVALID, "bad receiver list");
    std::vector<int64_t> offs(nrec);
    for (int r = 0; r < nrec; ++r) {
        if (nodes[r] < 0 || nodes[r] >= h->nn)
            return fail(PHC_ERR_INVALID, "receiver node %d outside 0..%lld", nodes[r], (long long)h->nn - 1);
        offs[r] = s2_off(h, nodes[r]);
    }
    PHC_HIP(hipStreamSynchronize(h->stream));
    s2_drop_graph(h);
    h->nrec = nrec;
    if (nrec > 0) {
        PHC_TRY(h->recoff.alloc(sizeof(int64_t) * nrec));
        PHC_HIP(hipMemcpy(h->recoff.ptr, offs.data(), sizeof(int64_t) * nrec, hipMemcpyHostToDevice));
    }
    return PHC_OK;
}

phc_status phc_sem2d_get_traces(phc_sem2d *h, float *src, float *rec)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    if (src && h->last_n > 0)
        PHC_HIP(hipMemcpy(src, h->trace.ptr, sizeof(float) * h->last_n, hipMemcpyDeviceToHost));
    if (rec && h->last_n > 0 && h->last_nrec > 0)
        PHC_HIP(hipMemcpy(rec, h->rec.ptr, sizeof(float) * (size_t)h->last_n * h->last_nrec,
                          hipMemcpyDeviceToHost));
    return PHC_OK;
}

phc_status phc_sem2d_get_field(phc_sem2d *h, int level, float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "NULL argument");
    if (level < 0 || level > 1) return fail(PHC_ERR_INVALID, "level %d is not 0 or 1", level);
    return s2_get(h, h->lev[level], host);
}

phc_status phc_sem2d_set_field(phc_sem2d *h, int level, const float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "NULL argument");
    if (level < 0 || level > 1) return fail(PHC_ERR_INVALID, "level %d is not 0 or 1", level);
    PHC_HIP(hipStreamSynchronize(h->stream));
    return s2_put(h, h->lev[level], host);
}

phc_status phc_sem2d_get_mass(phc_sem2d *h, float *mass, float *boundary)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (mass) std::memcpy(mass, h->hmass.data(), sizeof(float) * h->nn);
    if (boundary) std::memcpy(boundary, h->hsh.data(), sizeof(float) * h->nn);
    return PHC_OK;
}

phc_status phc_sem2d_check_finite(phc_sem2d *h, int level)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    std::vector<float> f((size_t)h->nn);
    PHC_TRY(phc_sem2d_get_field(h, level, f.data()));
    for (int64_t n = 0; n < h->nn; ++n)
        if (!std::isfinite(f[n])) return fail(PHC_ERR_NONFINITE, "level %d node %lld is not finite", level, (long long)n);
    return PHC_OK;
}

phc_status phc_sem2d_set_timing(phc_sem2d *h, int enable)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    h->timing = enable ? 1 : 0;
    return PHC_OK;
}

phc_status phc_sem2d_last_run_kernel_ms(phc_sem2d *h, double *ms)
{
    if (!h || !ms) return fail(PHC_ERR_INVALID, "NULL argument");
    *ms = h->last_ms;
    return PHC_OK;
}

phc_status phc_sem2d_synchronize(phc_sem2d *h)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

// Autotuner (no reference counterpart): rows per wave, timed on scratch
// copies of the levels (state untouched).
phc_status phc_sem2d_autotune(phc_sem2d *h, int steps_per_trial, int *chunk)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (steps_per_trial < 1) return fail(PHC_ERR_INVALID, "steps_per_trial must be >= 1");
    const size_t bytes = sizeof(float) * (size_t)h->L.total;
    DevBuf s0, s1;
    PHC_TRY(s0.alloc(bytes));
    phc_status st = s1.alloc(bytes);
    if (st != PHC_OK) { s0.release(); return st; }
    if (hipMemcpyAsync(s0.ptr, h->lev[0].ptr, bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(s1.ptr, h->lev[1].ptr, bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
        st = fail(PHC_ERR_HIP, "scratch copy failed");
    std::swap(h->lev[0], s0);
    std::swap(h->lev[1], s1);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st == PHC_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = fail(PHC_ERR_HIP, "event create failed");
    const int chunks[] = {1, 2, 4, 8, 16, 32, 64, 128, 256, 1 << 30};
    double best = 1e30;
    int bc = h->chunk;
    int last = -1;
    for (int c0 : chunks) {
        if (st != PHC_OK) break;
        const int c = std::min(c0, h->nz);
        if (c == last) continue;
        last = c;
        h->chunk = c;
        float ms = 0.f;
        for (int rep = 0; rep < 2 && st == PHC_OK; ++rep) {
            if (hipEventRecord(e0, h->stream) != hipSuccess) st = fail(PHC_ERR_HIP, "event record failed");
            for (int s = 0; s < steps_per_trial && st == PHC_OK; ++s)
                st = s2_launch(h, s & 1, 1 - (s & 1), 0, 0.f, nullptr, nullptr, nullptr);
            if (st == PHC_OK && (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                 hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                st = fail(PHC_ERR_HIP, "trial timing failed");
        }
        if (st == PHC_OK && ms < best) { best = ms; bc = c; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(h->stream);
    std::swap(h->lev[0], s0);
    std::swap(h->lev[1], s1);
    s0.release();
    s1.release();
    h->chunk = bc;
    s2_drop_graph(h);
    PHC_TRY(st);
    if (chunk) *chunk = h->chunk;
    return PHC_OK;
}

} // extern "C"
