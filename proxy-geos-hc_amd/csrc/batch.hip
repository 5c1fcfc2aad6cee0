// batch.hip -- many FD shots on uneven grids in one persistent launch per
// time step (phc_batch_* C ABI).
//
// No reference counterpart: the reference runs one shot per process (its FD
// driver, src/fd/FDTDmain.cpp:9-108).  Each shot here is exactly that FD
// problem -- same setup, same step (FDTDkernels.hpp:8-34, 108-128), same
// sponge -- on its own grid size, source and model.  One launch advances every
// shot by one step: the (shot, tile, chunk) work items of all shots are dealt
// to a fixed number of resident blocks, each block marching its items with the
// same tile body as the single-grid kernel (fd_tile).  The run is captured in
// one hipGraph; the autotuner picks block shape, chunk length and resident
// blocks per CU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "fd_kernels.hpp"
#include "fd_tile.hpp"
#include "xcd_sched.hpp"
#include "march.hpp"
#include "phc_common.hpp"

namespace phc {

// One shot: its layout and device arrays (inside the batch's pool).
struct ShotDesc {
    FdLayout L;
    float *lev[2];
    const float *vp;  // vp^2 dt^2
    const float *fy, *fz;
    const float *rhs;
    int64_t src_off;
    int krem;         // first k of the shot's remainder slab (0: none)
};

// One unit of work: tile column (kt, jt) of i-chunk ic of one shot; kt < 0:
// part jt of the shot's k-remainder slab, BATCH_SLAB_PASSES x the block's
// threads points in slab order.
// A tile column (k tile kt, j tile jt) of one shot marched over planes
// [i0, i1); kt < 0: part jt of the shot's remainder slab.
struct WorkItem {
    int shot, kt, jt, i0, i1;
};
constexpr int BATCH_SLAB_PASSES = 8;

// The per-grid fields fd_tile reads, for one work item (scalar registers).
struct FdGrid {
    FdLayout L;
    const float *pn;
    float *pnm1;
    const float *vp, *fy, *fz;
    int64_t src_off;
    int inj;
    float rhs_next;
    float *trace;
    int ilo, ihi, jlo, jhi, klo, khi;  // every point of the shot
    int krem;
    int jrem;  // no j-remainder slab in the batch (0)
};

// Work queues of one step: queue q holds items[start[q] .. start[q]+len[q]),
// a contiguous run of tiles in spatial order of roughly equal cost.
constexpr int BATCH_QUEUES = 8;  // one per XCD
struct BatchQueues {
    int start[BATCH_QUEUES];
    int len[BATCH_QUEUES];
};

// One time step of every shot.  `c` holds what the shots share (coefficients,
// sponge table, chunk length, deferred damping flag).  The hardware deals
// consecutive blocks round-robin over the 8 XCDs, so block b starts on queue
// b % 8: the blocks of one XCD take the tiles of one contiguous region in
// order, and neighbouring tiles that run side by side share their halo rows
// in that XCD's L2.  A block whose queue is empty takes items from the other
// queues.  counter[q] (zero at the start of the step) is the next unclaimed
// position of queue q.
template <int R, int TY, int NJ, int ARITH, typename IT>
__global__ void __launch_bounds__(FD_TK *TY)
fd_batch_kernel(FdStepArgs c, const ShotDesc *__restrict__ shots, const WorkItem *__restrict__ items,
                BatchQueues qs, int cur, int it, int inj, float *traces, int trace_stride, int step,
                int *counter)
{
    using SH = FdTileShape<R, TY, NJ>;
    __shared__ float tile[2][SH::LH][SH::LW];
    __shared__ int claimed[2];  // alternating slots: a slot is rewritten two claims later,
    int slot = 0;               // after a barrier every thread passes once it has read it
    const bool leader = threadIdx.x == 0 && threadIdx.y == 0;
    const int q0 = blockIdx.x % BATCH_QUEUES;
    for (int d = 0; d < BATCH_QUEUES; ++d) {
        const int q = (q0 + d) % BATCH_QUEUES;
        while (true) {
            if (leader) claimed[slot] = atomicAdd(counter + q, 1);
            lds_barrier();
            const int pos = claimed[slot];
            slot ^= 1;
            if (pos >= qs.len[q]) break;
            const WorkItem wi = items[qs.start[q] + pos];
            const ShotDesc &sd = shots[wi.shot];
            FdGrid g;
            g.L = sd.L;
            g.pn = sd.lev[cur];
            g.pnm1 = sd.lev[1 - cur];
            g.vp = sd.vp;
            g.fy = sd.fy;
            g.fz = sd.fz;
            g.src_off = sd.src_off;
            g.inj = inj;
            g.rhs_next = inj ? sd.rhs[it + 1] : 0.f;
            g.trace = traces + (int64_t)wi.shot * trace_stride + step;
            g.ilo = 0; g.ihi = sd.L.nx;
            g.jlo = 0; g.jhi = sd.L.ny;
            g.klo = 0; g.khi = sd.L.nz;
            g.krem = sd.krem;
            g.jrem = 0;
            if (wi.kt < 0) {
                constexpr int T = FD_TK * TY;
                for (int m = 0; m < BATCH_SLAB_PASSES; ++m)
                    fd_rem_points<R, ARITH, IT>(c, g, ((int64_t)wi.jt * BATCH_SLAB_PASSES + m) * T, T);
                continue;
            }
            fd_tile<R, TY, NJ, ARITH, IT>(c, g, wi.kt, wi.jt, wi.i0, wi.i1, tile);
        }
    }
}

// Zero the per-step work counters (a kernel, so that captured graphs replay it).
__global__ void fd_batch_zero_kernel(int *counter, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) counter[t] = 0;
}

// rhs[it] into every shot's current level at its source (FDTDkernels.hpp:121-128)
__global__ void fd_batch_inject_kernel(const ShotDesc *shots, int nshots, int cur, int it)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    const ShotDesc &d = shots[s];
    float *p = d.lev[cur];
    p[d.src_off] = add_nc(p[d.src_off], mul_nc(d.vp[d.src_off], d.rhs[it]));
}

// Receiver sampling of every shot after a step: out[(shot*nrec + r)*stride +
// step] = new level at the receiver.  A receiver on the source point reads the
// value recorded before the fused injection of the next step.
__global__ void fd_batch_sample_kernel(const ShotDesc *shots, int nshots, const int64_t *offs, int nrec,
                                       int level, float *out, int stride, int step, const float *traces)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nshots * nrec) return;
    const int s = t / nrec;
    const ShotDesc &d = shots[s];
    const int64_t o = offs[t];
    out[(int64_t)t * stride + step] = (o == d.src_off) ? traces[(int64_t)s * stride + step] : d.lev[level][o];
}

using BatchFn = void (*)(FdStepArgs, const ShotDesc *, const WorkItem *, BatchQueues, int, int, int, float *,
                        int, int, int *);

template <int R, int TY, int NJ>
static BatchFn batch_arith(int arith)
{
    return arith == PHC_FD_ARITH_REF ? fd_batch_kernel<R, TY, NJ, PHC_FD_ARITH_REF, uint32_t>
                                     : fd_batch_kernel<R, TY, NJ, PHC_FD_ARITH_FP32, uint32_t>;
}

static BatchFn batch_pick(int R, int ty, int nj, int arith)
{
    if (R != 4) return nullptr;
    if (ty == 16 && nj == 1) return batch_arith<4, 16, 1>(arith);
    if (ty == 8 && nj == 2) return batch_arith<4, 8, 2>(arith);
    if (ty == 4 && nj == 2) return batch_arith<4, 4, 2>(arith);
    if (ty == 4 && nj == 4) return batch_arith<4, 4, 4>(arith);
    return nullptr;
}

void fd_axis_damping(int n, int W, float alpha, std::vector<float> &f);

} // namespace phc

using namespace phc;

struct phc_batch {
    int nshots = 0, R = 4, nsamples = 0, arith = 0, sw = 0;
    float alpha = 0.f;
    float coef[FD_MAX_R + 1];
    double coef0 = 0.0;
    std::vector<FdLayout> L;
    std::vector<int> src;
    std::vector<ShotDesc> hdesc;
    DevBuf pool, shots, items, traces, rec, recoff, counters;
    int cur = 1;      // level holding the current field (the reference's i2)
    int pending = 0;  // the previous level still owes its sponge product
    hipStream_t stream = nullptr;
    int ty = 16, nj = 1, chunk_i = 64, blocks_per_cu = 1, cus = 256;
    int krem_mode = 0;  // phc_batch_params.k_remainder (as requested)
    int krem_eff = 0;   // the mode the work list uses (the autotuner may pick -1 for 0)
    int nitems = 0;
    BatchQueues queues;
    int nrec = 0;  // receivers per shot
    int timing = 0;
    double last_ms = 0.0;
    int last_n = 0, last_nrec = 0;
    hipGraphExec_t gexec = nullptr;
    int g_it0 = -1, g_n = -1, g_cur = -1, g_pending = -1, g_nrec = -1;
};

static void batch_drop_graph(phc_batch *h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    h->g_it0 = h->g_n = h->g_cur = h->g_pending = h->g_nrec = -1;
}

// Work items of every shot for the current shape, dealt to BATCH_QUEUES
// queues (one per XCD).  A shot whose last k tile would hold 1..32 points
// (mode 0; any partial tile with mode 1, none with -1) gets its remainder
// slab as point-by-point items instead.
// * chunk_i > 0: items in spatial order (shot, i-chunk, j-tile, k-tile,
//   then the slab parts), cut into contiguous runs of about equal cost
//   (planes incl. the 2R lead-in x rows; a slab point costs a few tile
//   points: it reads its stencil from memory).
// * chunk_i = -m: each shot's per-XCD schedule (xcd_sched.hpp, split m, one
//   round = the XCD's resident blocks): queue x takes the entries of XCD x,
//   shot after shot, so an XCD's blocks march one shot's plane range over a
//   band of its columns in phase, as the single-grid schedule does; slab
//   parts go to the queue with the least work so far.
static phc_status batch_build_items(phc_batch *h)
{
    std::vector<std::vector<WorkItem>> qitems(BATCH_QUEUES);
    std::vector<double> qcost(BATCH_QUEUES, 0.0);
    std::vector<WorkItem> items;
    std::vector<double> cost;
    const int TJ = h->ty * h->nj;
    const int T = FD_TK * h->ty;
    const bool sched = h->chunk_i < 0;
    const int slots = std::max(h->cus / BATCH_QUEUES, 1) * h->blocks_per_cu;
    std::vector<int4> tab;
    for (int s = 0; s < h->nshots; ++s) {
        const FdLayout &L = h->L[s];
        const int full = L.nz / FD_TK, rem = L.nz - full * FD_TK;
        const bool slab = h->krem_eff >= 0 && full > 0 && rem > 0 && (h->krem_eff == 1 || rem <= FD_TK / 2);
        h->hdesc[s].krem = slab ? FD_TK * full : 0;
        const int tk = slab ? full : (L.nz + FD_TK - 1) / FD_TK, tj = (L.ny + TJ - 1) / TJ;
        auto tile_cost = [&](int jt, int planes) {
            return (double)(planes + 2 * h->R) * std::min(TJ, L.ny - jt * TJ);
        };
        if (sched) {
            build_xcd_schedule(L.nx, 0, tk * tj, slots, -h->chunk_i, 1, tab);
            for (size_t b = 0; b < tab.size(); ++b) {
                const int4 e = tab[b];
                if (e.y >= e.z) continue;
                const int x = (int)(b % BATCH_QUEUES), kt = e.x % tk, jt = e.x / tk;
                qitems[x].push_back({s, kt, jt, e.y, e.z});
                qcost[x] += tile_cost(jt, e.z - e.y);
            }
        } else {
            const int nc = (L.nx + h->chunk_i - 1) / h->chunk_i;
            for (int ic = 0; ic < nc; ++ic)
                for (int jt = 0; jt < tj; ++jt)
                    for (int kt = 0; kt < tk; ++kt) {
                        const int i0 = ic * h->chunk_i, i1 = std::min(i0 + h->chunk_i, L.nx);
                        items.push_back({s, kt, jt, i0, i1});
                        cost.push_back(tile_cost(jt, i1 - i0));
                    }
        }
        if (slab) {
            const int64_t pts = (int64_t)rem * L.ny * L.nx, per = (int64_t)BATCH_SLAB_PASSES * T;
            for (int64_t q = 0; q * per < pts; ++q) {
                const double c = 3.0 * (double)std::min(per, pts - q * per) / FD_TK;
                if (sched) {
                    const int x = (int)(std::min_element(qcost.begin(), qcost.end()) - qcost.begin());
                    qitems[x].push_back({s, -1, (int)q, 0, 0});
                    qcost[x] += c;
                } else {
                    items.push_back({s, -1, (int)q, 0, 0});
                    cost.push_back(c);
                }
            }
        }
    }
    // the descriptors carry each shot's slab start
    PHC_HIP(hipMemcpy(h->shots.ptr, h->hdesc.data(), sizeof(ShotDesc) * h->nshots, hipMemcpyHostToDevice));
    if (sched) {
        for (int k = 0; k < BATCH_QUEUES; ++k) {
            h->queues.start[k] = (int)items.size();
            h->queues.len[k] = (int)qitems[k].size();
            items.insert(items.end(), qitems[k].begin(), qitems[k].end());
        }
    } else {
        const double total = std::accumulate(cost.begin(), cost.end(), 0.0);
        int q = 0;
        double acc = 0.0;
        h->queues.start[0] = 0;
        for (size_t i = 0; i < items.size(); ++i) {
            acc += cost[i];
            while (q + 1 < BATCH_QUEUES && acc >= total * (q + 1) / BATCH_QUEUES) h->queues.start[++q] = (int)i + 1;
        }
        while (q + 1 < BATCH_QUEUES) h->queues.start[++q] = (int)items.size();
        for (int k = 0; k < BATCH_QUEUES; ++k)
            h->queues.len[k] =
                (k + 1 < BATCH_QUEUES ? h->queues.start[k + 1] : (int)items.size()) - h->queues.start[k];
    }
    h->nitems = (int)items.size();
    PHC_TRY(h->items.alloc(sizeof(WorkItem) * items.size()));
    PHC_HIP(hipMemcpy(h->items.ptr, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice));
    return PHC_OK;
}

static void batch_common(const phc_batch *h, FdStepArgs &c, int damp)
{
    std::memset(&c, 0, sizeof(c));
    for (int l = 0; l <= FD_MAX_R; ++l) c.cx[l] = c.cy[l] = c.cz[l] = h->coef[l];
    c.coef0 = h->coef0;
    c.c0f = (float)h->coef0;
    c.sw = h->sw;
    for (int d = 0; d <= h->sw && d <= FD_MAX_SW; ++d)
        c.stab[d] = (float)std::exp((double)h->alpha * std::pow((double)d, 2.0));
    c.damp_prev = damp;
    c.chunk_i = h->chunk_i;
}

// One batched step: every shot from level `cur` (current) and 1-cur
// (previous, overwritten by the new level).  traces[shot*stride + step] gets
// each shot's new source value before the fused injection of rhs[it+1].
static phc_status batch_launch(phc_batch *h, int cur, int it, int inj, int damp, float *traces, int stride,
                               int step, int *counter)
{
    BatchFn fn = batch_pick(h->R, h->ty, h->nj, h->arith);
    if (!fn) return fail(PHC_ERR_INVALID, "no batched kernel for %d waves x %d rows", h->ty, h->nj);
    FdStepArgs c;
    batch_common(h, c, damp);
    // resident blocks, a multiple of the queue count (no more than there are items)
    const int grid = std::max(BATCH_QUEUES, std::min(h->nitems, h->cus * h->blocks_per_cu) / BATCH_QUEUES * BATCH_QUEUES);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(FD_TK, h->ty), 0, h->stream, c, h->shots.as<ShotDesc>(),
                       h->items.as<WorkItem>(), h->queues, cur, it, inj, traces, stride, step, counter);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

// The whole run on h->stream: injection of rhs[it0] into every current level,
// then one batched step per sample (plus receiver sampling).  With `ev`,
// events bracket every batched launch.
static phc_status batch_zero_counters(phc_batch *h, int n)
{
    hipLaunchKernelGGL(fd_batch_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, h->stream,
                       h->counters.as<int>(), n);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

static phc_status batch_enqueue(phc_batch *h, int it0, int nsteps, hipEvent_t *ev)
{
    PHC_TRY(batch_zero_counters(h, nsteps * BATCH_QUEUES));
    hipLaunchKernelGGL(fd_batch_inject_kernel, dim3((h->nshots + 63) / 64), dim3(64), 0, h->stream,
                       h->shots.as<ShotDesc>(), h->nshots, h->cur, it0);
    PHC_HIP(hipGetLastError());
    int cur = h->cur;
    for (int s = 0; s < nsteps; ++s) {
        const int damp = (s == 0) ? h->pending : (h->sw > 0 ? 1 : 0);
        if (ev) PHC_HIP(hipEventRecord(ev[2 * s], h->stream));
        PHC_TRY(batch_launch(h, cur, it0 + s, s + 1 < nsteps ? 1 : 0, damp, h->traces.as<float>(), nsteps, s,
                             h->counters.as<int>() + s * BATCH_QUEUES));
        if (ev) PHC_HIP(hipEventRecord(ev[2 * s + 1], h->stream));
        if (h->nrec > 0) {
            const int n = h->nshots * h->nrec;
            hipLaunchKernelGGL(fd_batch_sample_kernel, dim3((n + 255) / 256), dim3(256), 0, h->stream,
                               h->shots.as<ShotDesc>(), h->nshots, h->recoff.as<int64_t>(), h->nrec, 1 - cur,
                               h->rec.as<float>(), nsteps, s, h->traces.as<float>());
            PHC_HIP(hipGetLastError());
        }
        cur = 1 - cur;
    }
    return PHC_OK;
}

static phc_status batch_upload(const std::vector<float> &buf, float *dev)
{
    PHC_HIP(hipMemcpy(dev, buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice));
    return PHC_OK;
}

extern "C" {

phc_status phc_batch_create(const phc_batch_params *p, phc_batch **out)
{
    if (!out) return fail(PHC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!p) return fail(PHC_ERR_INVALID, "params is NULL");
    if (p->nshots < 1) return fail(PHC_ERR_INVALID, "no shots (nshots = %d)", p->nshots);
    if (!p->dims || !p->src || !p->vp || !p->rhs || !p->coef)
        return fail(PHC_ERR_INVALID, "dims, src, vp, rhs or coef is NULL");
    if (p->half_width != 4) return fail(PHC_ERR_INVALID, "batched shots need half width 4 (got %d)", p->half_width);
    if (p->sponge_width < 0 || p->sponge_width > FD_MAX_SW)
        return fail(PHC_ERR_INVALID, "sponge width %d outside 0..%d", p->sponge_width, FD_MAX_SW);
    if (p->arith != PHC_FD_ARITH_REF && p->arith != PHC_FD_ARITH_FP32)
        return fail(PHC_ERR_INVALID, "unknown arithmetic mode %d", p->arith);
    if (p->nsamples < 1) return fail(PHC_ERR_INVALID, "nsamples must be >= 1 (got %d)", p->nsamples);
    if (p->chunk_i < -4 || p->blocks_per_cu < 0 || p->blocks_per_cu > 8)
        return fail(PHC_ERR_INVALID, "bad chunk_i (planes per item, or -1..-4) or blocks_per_cu");
    if (p->k_remainder < -1 || p->k_remainder > 1)
        return fail(PHC_ERR_INVALID, "k_remainder must be -1, 0 or 1 (got %d)", p->k_remainder);
    const int ty = p->block_waves ? p->block_waves : 16;
    const int tj = p->tile_j ? p->tile_j : ty;
    if (ty <= 0 || tj % ty != 0 || !batch_pick(p->half_width, ty, tj / ty, p->arith))
        return fail(PHC_ERR_INVALID, "no batched kernel for tile_j %d, %d waves per block", tj, ty);
    for (int s = 0; s < p->nshots; ++s) {
        const int nx = p->dims[3 * s], ny = p->dims[3 * s + 1], nz = p->dims[3 * s + 2];
        const int xs = p->src[3 * s], ys = p->src[3 * s + 1], zs = p->src[3 * s + 2];
        if (nx < 1 || ny < 1 || nz < 1) return fail(PHC_ERR_INVALID, "shot %d: empty grid", s);
        if (xs < 0 || xs >= nx || ys < 0 || ys >= ny || zs < 0 || zs >= nz)
            return fail(PHC_ERR_INVALID, "shot %d: source (%d,%d,%d) outside the grid", s, xs, ys, zs);
        if (!p->vp[s] || !p->rhs[s]) return fail(PHC_ERR_INVALID, "shot %d: vp or rhs is NULL", s);
        if (make_fd_layout(nx, ny, nz, p->half_width).total * 4 >= ((int64_t)1 << 32))
            return fail(PHC_ERR_INVALID, "shot %d: arrays exceed 4 GiB", s);
    }
    phc_batch *h = new phc_batch();
    auto bail = [&](phc_status st) { phc_batch_destroy(h); return st; };
    h->nshots = p->nshots;
    h->R = p->half_width;
    h->nsamples = p->nsamples;
    h->arith = p->arith;
    h->sw = p->sponge_width;
    h->alpha = p->sponge_alpha;
    std::memset(h->coef, 0, sizeof(h->coef));
    for (int l = 0; l <= h->R; ++l) h->coef[l] = p->coef[l];
    h->coef0 = p->coef0;
    h->src.assign(p->src, p->src + 3 * h->nshots);
    h->ty = ty;
    h->nj = tj / ty;
    h->chunk_i = p->chunk_i ? p->chunk_i : 64;
    h->blocks_per_cu = p->blocks_per_cu ? p->blocks_per_cu : 1;
    h->krem_mode = h->krem_eff = p->k_remainder;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        h->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "hipStreamCreate failed"));
    // one pool: per shot two levels, vp^2dt^2, fy, fz, rhs; each shot 256-byte aligned
    std::vector<int64_t> off(h->nshots);
    int64_t total = 0;
    h->L.resize(h->nshots);
    for (int s = 0; s < h->nshots; ++s) {
        h->L[s] = make_fd_layout(p->dims[3 * s], p->dims[3 * s + 1], p->dims[3 * s + 2], h->R);
        off[s] = total;
        total += 3 * h->L[s].total + h->L[s].ny + h->L[s].nz + h->nsamples;
        total = round_up(total, 64);
    }
    phc_status st;
    if ((st = h->pool.alloc(sizeof(float) * total)) || (st = h->shots.alloc(sizeof(ShotDesc) * h->nshots)))
        return bail(st);
    h->hdesc.resize(h->nshots);
    for (int s = 0; s < h->nshots; ++s) {
        const FdLayout &L = h->L[s];
        float *base = h->pool.as<float>() + off[s];
        ShotDesc &d = h->hdesc[s];
        d.L = L;
        d.lev[0] = base;
        d.lev[1] = base + L.total;
        float *vp = base + 2 * L.total, *fy = vp + L.total, *fz = fy + L.ny, *rhs = fz + L.nz;
        d.vp = vp;
        d.fy = fy;
        d.fz = fz;
        d.rhs = rhs;
        d.src_off = fd_off(L, h->src[3 * s], h->src[3 * s + 1], h->src[3 * s + 2]);
        std::vector<float> buf((size_t)L.total, p->init_value);
        if ((st = batch_upload(buf, d.lev[0])) || (st = batch_upload(buf, d.lev[1]))) return bail(st);
        std::fill(buf.begin(), buf.end(), 0.f);
        const float *v = p->vp[s];
        for (int64_t n = 0; n < (int64_t)L.nx * L.ny * L.nz; ++n)
            if (!(v[n] >= 0.f) || !std::isfinite(v[n]))
                return bail(fail(PHC_ERR_INVALID, "shot %d: vp^2 dt^2 at %lld is negative or not finite", s,
                                 (long long)n));
        for (int i = 0; i < L.nx; ++i)
            for (int j = 0; j < L.ny; ++j)
                std::memcpy(&buf[(size_t)fd_off(L, i, j, 0)], v + ((int64_t)i * L.ny + j) * L.nz,
                            sizeof(float) * L.nz);
        if ((st = batch_upload(buf, vp))) return bail(st);
        std::vector<float> f;
        fd_axis_damping(L.ny, h->sw, h->alpha, f);
        if ((st = batch_upload(f, fy))) return bail(st);
        fd_axis_damping(L.nz, h->sw, h->alpha, f);
        if ((st = batch_upload(f, fz))) return bail(st);
        if ((st = batch_upload(std::vector<float>(p->rhs[s], p->rhs[s] + h->nsamples), rhs))) return bail(st);
    }
    if (hipMemcpy(h->shots.ptr, h->hdesc.data(), sizeof(ShotDesc) * h->nshots, hipMemcpyHostToDevice) !=
        hipSuccess)
        return bail(fail(PHC_ERR_HIP, "descriptor upload failed"));
    if ((st = batch_build_items(h))) return bail(st);
    *out = h;
    return PHC_OK;
}

phc_status phc_batch_destroy(phc_batch *h)
{
    if (!h) return PHC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    batch_drop_graph(h);
    for (DevBuf *b : {&h->pool, &h->shots, &h->items, &h->traces, &h->rec, &h->recoff, &h->counters})
        b->release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PHC_OK;
}

phc_status phc_batch_set_receivers(phc_batch *h, int nrec, const int *ijk)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nrec < 0 || (nrec > 0 && !ijk)) return fail(PHC_ERR_INVALID, "bad receiver list");
    std::vector<int64_t> offs((size_t)h->nshots * nrec);
    for (int s = 0; s < h->nshots; ++s)
        for (int r = 0; r < nrec; ++r) {
            const int *q = ijk + 3 * ((size_t)s * nrec + r);
            const FdLayout &L = h->L[s];
            if (q[0] < 0 || q[0] >= L.nx || q[1] < 0 || q[1] >= L.ny || q[2] < 0 || q[2] >= L.nz)
                return fail(PHC_ERR_INVALID, "shot %d receiver %d at (%d,%d,%d) is outside the grid", s, r, q[0],
                            q[1], q[2]);
            offs[(size_t)s * nrec + r] = fd_off(L, q[0], q[1], q[2]);
        }
    PHC_HIP(hipStreamSynchronize(h->stream));
    batch_drop_graph(h);
    h->nrec = nrec;
    if (nrec > 0) {
        PHC_TRY(h->recoff.alloc(sizeof(int64_t) * offs.size()));
        PHC_HIP(hipMemcpy(h->recoff.ptr, offs.data(), sizeof(int64_t) * offs.size(), hipMemcpyHostToDevice));
    }
    return PHC_OK;
}

phc_status phc_batch_run(phc_batch *h, int it0, int nsteps)
{
    RoctxRange rr_("phc_batch_run");
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nsteps < 0) return fail(PHC_ERR_INVALID, "negative step count");
    if (nsteps == 0) return PHC_OK;
    if (it0 < 0 || it0 + nsteps > h->nsamples)
        return fail(PHC_ERR_INVALID, "steps %d..%d outside the %d source samples", it0, it0 + nsteps - 1,
                    h->nsamples);
    const size_t tbytes = sizeof(float) * (size_t)nsteps * h->nshots;
    const size_t rbytes = tbytes * h->nrec;
    const size_t cbytes = sizeof(int) * (size_t)nsteps * BATCH_QUEUES;
    if (h->traces.bytes < tbytes || (h->nrec > 0 && h->rec.bytes < rbytes) || h->counters.bytes < cbytes) {
        PHC_HIP(hipStreamSynchronize(h->stream));
        batch_drop_graph(h);
        if (h->traces.bytes < tbytes) PHC_TRY(h->traces.alloc(tbytes));
        if (h->nrec > 0 && h->rec.bytes < rbytes) PHC_TRY(h->rec.alloc(rbytes));
        if (h->counters.bytes < cbytes) PHC_TRY(h->counters.alloc(cbytes));
    }
    if (h->timing) {
        std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
        for (auto &e : ev) PHC_HIP(hipEventCreate(&e));
        phc_status st = batch_enqueue(h, it0, nsteps, ev.data());
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
        if (!(h->gexec && h->g_it0 == it0 && h->g_n == nsteps && h->g_cur == h->cur &&
              h->g_pending == h->pending && h->g_nrec == h->nrec)) {
            batch_drop_graph(h);
            hipGraph_t graph = nullptr;
            PHC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            phc_status st = batch_enqueue(h, it0, nsteps, nullptr);
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
            h->g_it0 = it0; h->g_n = nsteps; h->g_cur = h->cur; h->g_pending = h->pending; h->g_nrec = h->nrec;
        }
        PHC_HIP(hipGraphLaunch(h->gexec, h->stream));
    }
    if (nsteps % 2) h->cur = 1 - h->cur;
    h->pending = h->sw > 0 ? 1 : 0;
    h->last_n = nsteps;
    h->last_nrec = h->nrec;
    return PHC_OK;
}

phc_status phc_batch_get_traces(phc_batch *h, float *src, float *rec)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->last_n * h->nshots;
    if (src && n > 0) PHC_HIP(hipMemcpy(src, h->traces.ptr, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (rec && n > 0 && h->last_nrec > 0)
        PHC_HIP(hipMemcpy(rec, h->rec.ptr, sizeof(float) * n * h->last_nrec, hipMemcpyDeviceToHost));
    return PHC_OK;
}

phc_status phc_batch_get_field(phc_batch *h, int shot, int current, float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "handle or host buffer is NULL");
    if (shot < 0 || shot >= h->nshots) return fail(PHC_ERR_INVALID, "shot %d outside 0..%d", shot, h->nshots - 1);
    PHC_HIP(hipStreamSynchronize(h->stream));
    const FdLayout &L = h->L[shot];
    const int level = current ? h->cur : 1 - h->cur;
    std::vector<float> buf((size_t)L.total);
    PHC_HIP(hipMemcpy(buf.data(), h->hdesc[shot].lev[level], sizeof(float) * L.total, hipMemcpyDeviceToHost));
    const int R = L.R, NYe = L.ny + 2 * R, NZe = L.nz + 2 * R, W = h->sw;
    const bool damp = !current && h->pending;
    std::vector<float> fx, fy, fz;
    fd_axis_damping(L.nx, W, h->alpha, fx);
    fd_axis_damping(L.ny, W, h->alpha, fy);
    fd_axis_damping(L.nz, W, h->alpha, fz);
    for (int i = -R; i < L.nx + R; ++i)
        for (int j = -R; j < L.ny + R; ++j) {
            float *dst = host + ((int64_t)(i + R) * NYe + (j + R)) * NZe;
            std::memcpy(dst, &buf[(size_t)fd_off(L, i, j, -R)], sizeof(float) * NZe);
            if (!damp || i < 0 || i >= L.nx || j < 0 || j >= L.ny) continue;
            const bool inI = i >= W && i < L.nx - W, inJ = j >= W && j < L.ny - W;
            for (int k = 0; k < L.nz; ++k) {
                const bool inK = k >= W && k < L.nz - W;
                float s = 1.f;
                if (inJ && inK && !inI) s = fx[i];
                if (inK && inI && !inJ) s = fy[j];
                if (inJ && inI && !inK) s = fz[k];
                dst[k + R] = dst[k + R] * s;
            }
        }
    return PHC_OK;
}

phc_status phc_batch_set_timing(phc_batch *h, int enable)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    h->timing = enable ? 1 : 0;
    return PHC_OK;
}

phc_status phc_batch_last_run_kernel_ms(phc_batch *h, double *ms)
{
    if (!h || !ms) return fail(PHC_ERR_INVALID, "handle or ms is NULL");
    *ms = h->last_ms;
    return PHC_OK;
}

phc_status phc_batch_synchronize(phc_batch *h)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

phc_status phc_batch_autotune(phc_batch *h, int steps_per_trial, int *tile_j, int *block_waves, int *chunk_i,
                              int *blocks_per_cu)
{
    RoctxRange rr_("phc_batch_autotune");
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (steps_per_trial < 1 || steps_per_trial > 1000)
        return fail(PHC_ERR_INVALID, "steps_per_trial %d outside 1..1000", steps_per_trial);
    PHC_HIP(hipStreamSynchronize(h->stream));
    batch_drop_graph(h);
    // trials run on a copy of the pool; the shots' state is restored after
    DevBuf save, tr, ctr;
    PHC_TRY(save.alloc(h->pool.bytes));
    PHC_TRY(tr.alloc(sizeof(float) * h->nshots));
    const int nctr = steps_per_trial * BATCH_QUEUES;
    PHC_TRY(ctr.alloc(sizeof(int) * nctr));
    PHC_HIP(hipMemcpy(save.ptr, h->pool.ptr, h->pool.bytes, hipMemcpyDeviceToDevice));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PHC_HIP(hipEventCreate(&e0));
    PHC_HIP(hipEventCreate(&e1));
    static const int shapes[][2] = {{16, 1}, {8, 2}, {4, 2}, {4, 4}};
    static const int chunks[] = {32, 64, 128, 256, 512, -1, -2, -3, -4};
    static const int bpcs[] = {1, 2, 4};
    double best = 1e30;
    int bty = h->ty, bnj = h->nj, bci = h->chunk_i, bbp = h->blocks_per_cu, bku = h->krem_eff;
    phc_status st = PHC_OK;
    // requested mode 0: also the partial k tiles marched as tiles (-1)
    const int modes[2] = {h->krem_mode, -1};
    for (int mi = 0; mi < (h->krem_mode == 0 ? 2 : 1); ++mi)
    for (const auto &sh : shapes)
        for (int ci : chunks)
            for (int bp : bpcs) {
                if (st != PHC_OK) break;
                if (sh[0] * bp > 32) continue;  // at most 32 waves of one shape per CU
                h->ty = sh[0];
                h->nj = sh[1];
                h->chunk_i = ci;
                h->blocks_per_cu = bp;
                h->krem_eff = modes[mi];
                if ((st = batch_build_items(h))) break;
                float ms = 0.f;
                for (int rep = 0; rep < 2 && st == PHC_OK; ++rep) {  // first pass warms up
                    hipLaunchKernelGGL(fd_batch_zero_kernel, dim3((nctr + 255) / 256), dim3(256), 0, h->stream,
                                       ctr.as<int>(), nctr);
                    if (hipEventRecord(e0, h->stream) != hipSuccess) st = fail(PHC_ERR_HIP, "event record failed");
                    for (int s = 0; s < steps_per_trial && st == PHC_OK; ++s)
                        st = batch_launch(h, (h->cur + s) & 1, 0, 0, h->sw > 0, tr.as<float>(), 1, 0,
                                          ctr.as<int>() + s * BATCH_QUEUES);
                    if (st == PHC_OK && (hipEventRecord(e1, h->stream) != hipSuccess ||
                                         hipEventSynchronize(e1) != hipSuccess ||
                                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                        st = fail(PHC_ERR_HIP, "trial timing failed");
                }
                if (st == PHC_OK && ms < best) {
                    best = ms;
                    bty = h->ty; bnj = h->nj; bci = h->chunk_i; bbp = h->blocks_per_cu; bku = h->krem_eff;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(h->stream);
    if (hipMemcpy(h->pool.ptr, save.ptr, h->pool.bytes, hipMemcpyDeviceToDevice) != hipSuccess && st == PHC_OK)
        st = fail(PHC_ERR_HIP, "restoring the shots after tuning failed");
    save.release();
    tr.release();
    ctr.release();
    h->ty = bty; h->nj = bnj; h->chunk_i = bci; h->blocks_per_cu = bbp; h->krem_eff = bku;
    PHC_TRY(st);
    PHC_TRY(batch_build_items(h));
    if (tile_j) *tile_j = bty * bnj;
    if (block_waves) *block_waves = bty;
    if (chunk_i) *chunk_i = bci;
    if (blocks_per_cu) *blocks_per_cu = bbp;
    return PHC_OK;
}

} // extern "C"
