// mlp.hip -- learned thin absorbing boundary: training-data collection from
// wide-sponge runs, training (forward, backward, Adam) and MFMA inference
// fused into the FD time loop (phc_lb_* C ABI).  The model is described in
// lb.hpp.  No reference counterpart (the reference's boundaries are the
// sponge, src/fd/utilsFDTD/FDTDinit.hpp:319-396, and the PML,
// FDTDkernels.hpp:36-104).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "fd_kernels.hpp"
#include "lb.hpp"
#include "march.hpp"
#include "phc_common.hpp"

namespace phc {

// ------------------------------------------------------------ bf16 helpers
__device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------ layer geometry
// Face f = 2*axis + side.  Its layer is LB_W x n_u x n_v points (u, v the
// other two axes in increasing order); compact index ((d*n_u)+u)*n_v+v.
struct LbFaceGeom {
    int axis, side, nu, nv;
};

__host__ __device__ inline LbFaceGeom lb_face(const int n[3], int f)
{
    LbFaceGeom g;
    g.axis = f >> 1;
    g.side = f & 1;
    g.nu = (g.axis == 0) ? n[1] : n[0];
    g.nv = (g.axis == 2) ? n[1] : n[2];
    return g;
}

// grid coordinates of compact index (d, u, v) of face g
__host__ __device__ inline void lb_point(const int n[3], const LbFaceGeom &g, int d, int u, int v, int c[3])
{
    const int a = g.axis;
    const int ua = (a == 0) ? 1 : 0, va = (a == 2) ? 1 : 2;
    c[a] = g.side ? n[a] - 1 - d : d;
    c[ua] = u;
    c[va] = v;
}

int64_t lb_corr_layout(const FdLayout &L, int64_t *off)
{
    const int n[3] = {L.nx, L.ny, L.nz};
    off[0] = 0;
    for (int f = 0; f < LB_FACES; ++f) {
        const LbFaceGeom g = lb_face(n, f);
        off[f + 1] = off[f] + (int64_t)LB_W * g.nu * g.nv;
    }
    return off[LB_FACES];
}

__device__ inline uint32_t lb_hash(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (uint32_t)x;
}

// Inputs of one layer point from a field pair: x[0..8] = (pn - h)/s at
// offsets -4..4 along the inward normal, x[9..17] the same for pnm1, x[18] =
// sqrt(vp2 c1).  Offsets beyond the edge of the region [0, n) read as h (as
// the grid's halo holds h).  Returns s (0 if every input equals h).
struct LbGridView {
    const float *pn, *pnm1, *vp;
    FdLayout L;
    int o[3];  // region origin in the grid
    int n[3];  // region size
};

// The loads of lb_inputs: raw pn / pnm1 along the normal (x[0..17]) and vp at
// the point (x[18]); every offset is read (beyond the region's edge the grid
// still holds values: its halo, or the wide grid around the region), so the
// loads are not inside lane-divergent branches.
__device__ inline void lb_load_raw(const LbGridView &G, const LbFaceGeom &g, int d, int u, int v,
                                   float x[LB_NIN], int64_t *center)
{
    int c[3];
    lb_point(G.n, g, d, u, v, c);
    const int dir = g.side ? -1 : 1;  // inward
    int64_t step;
    if (g.axis == 0) step = G.L.pitch_j;
    else if (g.axis == 1) step = G.L.pitch_k;
    else step = 1;
    const int64_t base = fd_off(G.L, G.o[0] + c[0], G.o[1] + c[1], G.o[2] + c[2]);
    *center = base;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        const int64_t o = base + (int64_t)(m - 4) * dir * step;
        x[m] = G.pn[o];
        x[9 + m] = G.pnm1[o];
    }
    x[18] = G.vp[base];
}

// The arithmetic of lb_inputs on raw values from lb_load_raw (in place).
__device__ inline float lb_finish_inputs(int d, float h, float c1, float x[LB_NIN])
{
    float s = 0.f;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        const bool inside = 4 - m <= d;  // outward offset 4-m within the region
        const float a = inside ? x[m] - h : 0.f, b = inside ? x[9 + m] - h : 0.f;
        x[m] = a;
        x[9 + m] = b;
        s = fmaxf(s, fmaxf(fabsf(a), fabsf(b)));
    }
    // below LB_S_TINY the point is treated as quiet (1/s would overflow)
    const float inv = s > LB_S_TINY ? 1.f / s : 0.f;
#pragma unroll
    for (int m = 0; m < 18; ++m) x[m] *= inv;
    x[18] = sqrtf(x[18] * c1);
    return s;
}

__device__ inline float lb_inputs(const LbGridView &G, const LbFaceGeom &g, int d, int u, int v, float h, float c1,
                                  float x[LB_NIN], int64_t *center)
{
    lb_load_raw(G, g, d, u, v, x, center);
    return lb_finish_inputs(d, h, c1, x);
}

// ------------------------------------------------------------ collection
struct LbCollectArgs {
    LbGridView G;
    float h, c1, chat[LB_R + 1];
    float keep;   // probability of keeping a point (hash-based)
    float s_min;  // points whose s is below this are skipped
    uint64_t seed;
    float *out[LB_FACES];
    unsigned long long *count;  // per face
    int64_t cap;
    int64_t face_off[LB_FACES + 1];  // compact offsets of the region's faces
};

// Ordered append: every thread of the block calls this with its keep flag;
// kept threads get consecutive slots from *base in thread order (wave ballots
// and a prefix over the block's waves), and *base advances by the block's
// total.  Sample stores therefore fill in the same order on every run.
constexpr int LB_CB = 256;  // threads per collection block
__device__ inline unsigned long long lb_ordered_slot(bool kept, unsigned long long *base, int *wtot)
{
    const unsigned long long bal = __ballot(kept);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) wtot[wv] = __popcll(bal);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < LB_CB / 64; ++w) {
        before += (w < wv) ? wtot[w] : 0;
        total += wtot[w];
    }
    const unsigned long long slot = *base + (unsigned long long)before +
                                    (unsigned long long)__popcll(bal & ((1ull << lane) - 1ull));
    __syncthreads();
    if (threadIdx.x == 0) *base += (unsigned long long)total;
    __syncthreads();
    return slot;
}

// One block per face, walking its layer points in compact order (d, u, v).
// Target: the term the region's update misses,
// t = sum_{l>d} (c_l/c_1) (p(P + l n_out) - h) / s, read from the wide grid
// around the region.
__global__ void __launch_bounds__(LB_CB) lb_collect_kernel(LbCollectArgs A)
{
    __shared__ int wtot[LB_CB / 64];
    __shared__ unsigned long long base;
    const int f = blockIdx.x;
    const LbFaceGeom g = lb_face(A.G.n, f);
    const int64_t nf = A.face_off[f + 1] - A.face_off[f];
    if (threadIdx.x == 0) base = A.count[f];
    __syncthreads();
    const int dir = g.side ? -1 : 1;
    const int64_t step = g.axis == 0 ? A.G.L.pitch_j : (g.axis == 1 ? A.G.L.pitch_k : 1);
    for (int64_t c0 = 0; c0 < nf; c0 += LB_CB) {
        const int64_t r0 = c0 + threadIdx.x;
        bool kept = false;
        float x[LB_NIN];
        float tgt = 0.f;
        int d = 0;
        if (r0 < nf) {
            const int64_t t = A.face_off[f] + r0;
            int64_t r = r0;
            const int v = (int)(r % g.nv);
            r /= g.nv;
            const int u = (int)(r % g.nu);
            d = (int)(r / g.nu);
            const uint32_t hv = lb_hash(A.seed * 0x9e3779b97f4a7c15ull + (uint64_t)t);
            if ((float)(hv >> 8) * (1.f / 16777216.f) < A.keep) {
                int64_t center;
                const float sc = lb_inputs(A.G, g, d, u, v, A.h, A.c1, x, &center);
                if (sc >= A.s_min && sc > LB_S_TINY) {
                    for (int l = d + 1; l <= LB_R; ++l)
                        tgt += A.chat[l] * (A.G.pn[center - (int64_t)l * dir * step] - A.h);
                    tgt /= sc;
                    kept = true;
                }
            }
        }
        const unsigned long long slot = lb_ordered_slot(kept, &base, wtot);
        if (kept && (int64_t)slot < A.cap) {
            float *o = A.out[f] + slot * LB_NSTRIDE;
            for (int m = 0; m < LB_NIN; ++m) o[m] = x[m];
            o[LB_NIN] = tgt;
            o[LB_NIN + 1] = (float)d;
            o[LB_NIN + 2] = (float)r0;  // compact index in the face's layer
        }
    }
    if (threadIdx.x == 0) A.count[f] = base;
}

// On-policy samples: a thin run (the region itself, with the learned layer)
// and a wide run (the region plus margin and sponge) of the same shot at the
// same step n.  Inputs come from the thin run's own levels (the states the
// network meets when it runs), the target from the wide run's field beyond
// the region: t = sum_{l>d} (c_l/c_1) (W^n(P + l n_out) - h) / s_thin, the
// term a transparent edge would add at the thin run's own amplitude scale.
struct LbOnPolicyArgs {
    LbGridView T;              // thin levels n (pn) and n-1 (pnm1), o = 0, n = thin size
    const float *wcur;         // wide level n
    FdLayout LW;
    int wo[3];                 // region origin in the wide grid
    float h, c1;
    float chat[LB_R + 1];
    float keep, s_min, t_max;
    uint64_t seed;
    float *out[LB_FACES];
    unsigned long long *count;
    int64_t cap;
    int64_t face_off[LB_FACES + 1];
};

// One block per face of the thin run, layer points in compact order.
__global__ void __launch_bounds__(LB_CB) lb_onpolicy_kernel(LbOnPolicyArgs A)
{
    __shared__ int wtot[LB_CB / 64];
    __shared__ unsigned long long base;
    const int f = blockIdx.x;
    const LbFaceGeom g = lb_face(A.T.n, f);
    const int64_t nf = A.face_off[f + 1] - A.face_off[f];
    if (threadIdx.x == 0) base = A.count[f];
    __syncthreads();
    const int dir = g.side ? -1 : 1;
    const int64_t step = g.axis == 0 ? A.LW.pitch_j : (g.axis == 1 ? A.LW.pitch_k : 1);
    for (int64_t c0 = 0; c0 < nf; c0 += LB_CB) {
        const int64_t r0 = c0 + threadIdx.x;
        bool kept = false;
        float x[LB_NIN];
        float tgt = 0.f;
        int d = 0;
        if (r0 < nf) {
            const int64_t t = A.face_off[f] + r0;
            int64_t r = r0;
            const int v = (int)(r % g.nv);
            r /= g.nv;
            const int u = (int)(r % g.nu);
            d = (int)(r / g.nu);
            const uint32_t hv = lb_hash(A.seed * 0x9e3779b97f4a7c15ull + (uint64_t)t);
            if ((float)(hv >> 8) * (1.f / 16777216.f) < A.keep) {
                int64_t P;
                const float sc = lb_inputs(A.T, g, d, u, v, A.h, A.c1, x, &P);
                if (sc >= A.s_min && sc > LB_S_TINY) {
                    int c[3];
                    lb_point(A.T.n, g, d, u, v, c);
                    const int64_t Pw = fd_off(A.LW, A.wo[0] + c[0], A.wo[1] + c[1], A.wo[2] + c[2]);
                    for (int l = d + 1; l <= LB_R; ++l)
                        tgt += A.chat[l] * (A.wcur[Pw - (int64_t)l * dir * step] - A.h);
                    tgt /= sc;
                    // the thin state has lost the wave: no usable label
                    kept = fabsf(tgt) <= A.t_max;
                }
            }
        }
        const unsigned long long slot = lb_ordered_slot(kept, &base, wtot);
        if (kept && (int64_t)slot < A.cap) {
            float *o = A.out[f] + slot * LB_NSTRIDE;
            for (int m = 0; m < LB_NIN; ++m) o[m] = x[m];
            o[LB_NIN] = tgt;
            o[LB_NIN + 1] = (float)d;
            o[LB_NIN + 2] = (float)r0;
        }
    }
    if (threadIdx.x == 0) A.count[f] = base;
}

// ------------------------------------------------------------ training
constexpr int LB_TB = 64;  // samples (threads) per training block

struct LbBatch {
    const float *samples[LB_FACES];
    const int *idx[LB_FACES];  // positions in samples (nullptr: 0..count-1)
    int count[LB_FACES];
    int block_start[LB_FACES + 1];
    float inv_total;  // 1 / total samples in the batch
    float noise;      // std of the Gaussian noise added to the field inputs (0: none)
    uint64_t nseed;
    // non-null: each block writes its gradient sum to partial[block][p] and
    // its loss to ploss[block] (summed in a fixed order by lb_reduce_kernel,
    // so training repeats bit for bit); null: float atomics into grad, loss
    float *partial, *ploss;
};

// Standard normal from four hashed uniforms (Irwin-Hall, rescaled to unit variance).
__device__ inline float lb_gauss(uint64_t key)
{
    float acc = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc += (float)(lb_hash(key * 4ull + (uint64_t)r) >> 8) * (1.f / 16777216.f);
    return (acc - 2.f) * 1.7320508f;
}

// Forward, loss and backward of one batch: every block takes LB_TB samples of
// one face; the weights are used rounded to bf16 (the stored inference
// precision), arithmetic in FP32.  Per-sample activations and deltas go to
// LDS; each thread then sums the gradient of a strided subset of the face's
// parameters over the block's samples and adds it to grad (FP32 atomics).
__global__ void __launch_bounds__(LB_TB) lb_train_kernel(const float *__restrict__ theta, float *grad, float *loss,
                                                         LbBatch B)
{
    __shared__ float w[LB_NP];
    __shared__ float sx[LB_TB][LB_NIN + 1];
    __shared__ float sh1[LB_TB][LB_H + 1], sh2[LB_TB][LB_H + 1], sd1[LB_TB][LB_H + 1], sd2[LB_TB][LB_H + 1];
    __shared__ float sdy[LB_TB];
    __shared__ int sdep[LB_TB];
    __shared__ float sloss[LB_TB];
    int f = 0;
    while (f + 1 < LB_FACES && (int)blockIdx.x >= B.block_start[f + 1]) ++f;
    const int chunk = blockIdx.x - B.block_start[f];
    const int tid = threadIdx.x;
    for (int p = tid; p < LB_NP; p += LB_TB) w[p] = bf2f(f2bf(theta[f * LB_NP + p]));
    __syncthreads();
    const int pos = chunk * LB_TB + tid;
    const bool valid = pos < B.count[f];
    float t = 0.f;
    int d = 0;
    if (valid) {
        const int si = B.idx[f] ? B.idx[f][pos] : pos;
        const float *s = B.samples[f] + (int64_t)si * LB_NSTRIDE;
        for (int m = 0; m < LB_NIN; ++m) sx[tid][m] = s[m];
        t = s[LB_NIN];
        d = min(max((int)s[LB_NIN + 1], 0), LB_W - 1);
        if (B.noise > 0.f) {
            // denoising: perturb the field inputs that exist at this depth
            // (offsets inside the region), keep the clean target, so the
            // network learns not to amplify errors in its own inputs
            const uint64_t key = (B.nseed * 0x9e3779b97f4a7c15ull) ^ ((uint64_t)f << 40) ^ (uint64_t)si;
            for (int m = 0; m < 18; ++m)
                if (m % 9 >= 4 - d) sx[tid][m] += B.noise * lb_gauss(key * 32ull + (uint64_t)m);
        }
    } else {
        for (int m = 0; m < LB_NIN; ++m) sx[tid][m] = 0.f;
    }
    // forward; activations stay in this thread's LDS rows
    for (int i = 0; i < LB_H; ++i) {
        float a = w[LB_OFF_B1 + d * LB_H + i];
#pragma unroll
        for (int m = 0; m < LB_NIN; ++m) a = fmaf(w[LB_OFF_W1 + i * LB_NIN + m], sx[tid][m], a);
        sh1[tid][i] = fmaxf(a, 0.f);
    }
    float y = w[LB_OFF_B3];
    for (int i = 0; i < LB_H; ++i) {
        float a = w[LB_OFF_B2 + i];
#pragma unroll 8
        for (int j = 0; j < LB_H; ++j) a = fmaf(w[LB_OFF_W2 + i * LB_H + j], sh1[tid][j], a);
        const float h2 = fmaxf(a, 0.f);
        sh2[tid][i] = h2;
        y = fmaf(w[LB_OFF_W3 + i], h2, y);
    }
    // backward
    const float e = valid ? y - t : 0.f;
    const float dy = 2.f * B.inv_total * e;
    sloss[tid] = e * e;
    sdy[tid] = dy;
    sdep[tid] = d;
    for (int i = 0; i < LB_H; ++i) sd2[tid][i] = sh2[tid][i] > 0.f ? dy * w[LB_OFF_W3 + i] : 0.f;
    for (int j = 0; j < LB_H; ++j) {  // W2^T d2, masked by relu'
        float a = 0.f;
#pragma unroll 8
        for (int i = 0; i < LB_H; ++i) a = fmaf(w[LB_OFF_W2 + i * LB_H + j], sd2[tid][i], a);
        sd1[tid][j] = sh1[tid][j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid == 0) {
        float sum = 0.f;
        for (int k = 0; k < LB_TB; ++k) sum += sloss[k];
        if (B.ploss) B.ploss[blockIdx.x] = sum * B.inv_total;
        else atomicAdd(loss, sum * B.inv_total);
    }
    float *gf = grad + f * LB_NP;
    for (int p = tid; p < LB_NP; p += LB_TB) {
        float g = 0.f;
        if (p < LB_OFF_B1) {
            const int i = p / LB_NIN, m = p % LB_NIN;
            for (int k = 0; k < LB_TB; ++k) g = fmaf(sd1[k][i], sx[k][m], g);
        } else if (p < LB_OFF_W2) {
            const int dd = (p - LB_OFF_B1) / LB_H, i = (p - LB_OFF_B1) % LB_H;
            for (int k = 0; k < LB_TB; ++k) g += (sdep[k] == dd) ? sd1[k][i] : 0.f;
        } else if (p < LB_OFF_B2) {
            const int i = (p - LB_OFF_W2) / LB_H, j = (p - LB_OFF_W2) % LB_H;
            for (int k = 0; k < LB_TB; ++k) g = fmaf(sd2[k][i], sh1[k][j], g);
        } else if (p < LB_OFF_W3) {
            const int i = p - LB_OFF_B2;
            for (int k = 0; k < LB_TB; ++k) g += sd2[k][i];
        } else if (p < LB_OFF_B3) {
            const int i = p - LB_OFF_W3;
            for (int k = 0; k < LB_TB; ++k) g = fmaf(sdy[k], sh2[k][i], g);
        } else {
            for (int k = 0; k < LB_TB; ++k) g += sdy[k];
        }
        if (B.partial) B.partial[(int64_t)blockIdx.x * LB_NP + p] = g;
        else atomicAdd(gf + p, g);
    }
}

// Gradient of each face = its blocks' partial sums added in block order;
// loss = the blocks' partial losses added in block order.
__global__ void lb_reduce_kernel(LbBatch B, float *grad, float *loss)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < LB_FACES * LB_NP) {
        const int f = t / LB_NP, p = t - f * LB_NP;
        float g = 0.f;
        for (int b = B.block_start[f]; b < B.block_start[f + 1]; ++b) g += B.partial[(int64_t)b * LB_NP + p];
        grad[t] = g;
    }
    if (t == 0) {
        float acc = 0.f;
        for (int b = 0; b < B.block_start[LB_FACES]; ++b) acc += B.ploss[b];
        *loss = acc;
    }
}

// Random minibatch positions: idx[f][n] uniform in [0, count_f).
__global__ void lb_random_idx_kernel(int *idx, int n, int count, uint64_t seed)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) idx[t] = (int)(lb_hash(seed + 0x632be59bd9b4e019ull * (uint64_t)(t + 1)) % (uint32_t)count);
}

// Adam step on every parameter; clears the gradient.
__global__ void lb_adam_kernel(float *theta, float *grad, float *m, float *v, int n, float lr, float b1, float b2,
                               float c1, float c2, float eps)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float g = grad[t];
    const float mt = b1 * m[t] + (1.f - b1) * g;
    const float vt = b2 * v[t] + (1.f - b2) * g * g;
    m[t] = mt;
    v[t] = vt;
    theta[t] -= lr * (mt * c1) / (sqrtf(vt * c2) + eps);
    grad[t] = 0.f;
}

__global__ void lb_fill_kernel(float *p, int n, float val)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = val;
}

__global__ void lb_to_bf16_kernel(const float *theta, uint16_t *w, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) w[t] = f2bf(theta[t]);
}

// ------------------------------------------------------------ inference
// MFMA 32x32x16 (bf16 in, FP32 accumulate; twice the products per cycle of
// 32x32x8 on gfx950), one wavefront per 32 points of one face.  The network
// is evaluated transposed, H^T = W x^T, so that lane l always holds point
// l%32: operand A = a weight matrix (lane l: row l%32, k = 8*(l/32)+0..7 of
// the 16-wide k-chunk), operand B = activations (lane l: column l%32, the
// same k), result (lane l: column l%32, rows 8*(v/4) + 4*(l/32) + v%4 for
// v = 0..15).
//   Layer 1: chunk 0 holds the 9 field inputs' first 8 (half-wave 0: the
// current level's, half-wave 1: the previous level's), chunk 1 the ninth and
// (half-wave 0) the velocity feature; the weight columns are placed to match,
// so each lane feeds only its own level's inputs.
//   Layer 2: the 16 rows a lane holds are hidden units; chunk c takes the
// lane's results v = 8c..8c+7, and W2's columns are loaded in that order, so
// activations never leave the lane.
//   Biases enter as the accumulator operand of the first product of a layer.
// Activations (inputs and hidden) are split into a bf16 high part and a bf16
// remainder, two products each, so they carry ~16 bits.
typedef __bf16 lb_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lb_f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 lb_bf16x2_t __attribute__((ext_vector_type(2)));
typedef float lb_f32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t lb_u32x4 __attribute__((ext_vector_type(4)));

// Two floats to packed bf16, round to nearest even (one v_cvt_pk_bf16_f32;
// the same values as f2bf for finite inputs), element 0 in the low half.
__device__ inline uint32_t f2bf2(float a, float b)
{
    const lb_f32x2_t v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, lb_bf16x2_t));
}

// a = hi + lo with hi = bf16(a) and lo = bf16(a - hi): ~16 bits of a.
__device__ inline void lb_split8(const float a[8], lb_bf16x8 &hi, lb_bf16x8 &lo)
{
    lb_u32x4 h, l;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        h[j] = f2bf2(a[2 * j], a[2 * j + 1]);
        l[j] = f2bf2(a[2 * j] - __uint_as_float(h[j] << 16), a[2 * j + 1] - __uint_as_float(h[j] & 0xffff0000u));
    }
    hi = __builtin_bit_cast(lb_bf16x8, h);
    lo = __builtin_bit_cast(lb_bf16x8, l);
}

// Input of layer-1 slot (chunk c, half-wave h, element e), or -1 (zero weight).
__host__ __device__ constexpr int lb_in_slot(int c, int h, int e)
{
    return c == 0 ? (h == 0 ? e : 9 + e) : (e == 0 ? (h == 0 ? 8 : 17) : (e == 1 && h == 0 ? LB_NIN - 1 : -1));
}

// Hidden unit held by result register v of half-wave h.
__host__ __device__ constexpr int lb_row(int v, int h) { return 8 * (v >> 2) + 4 * h + (v & 3); }

// One face's weights in the registers of a wavefront, in MFMA operand layout
// (A operands of layers 1 and 2) or per result register (b2, w3).
struct LbWaveWeights {
    lb_bf16x8 a1[2], a2[2];
    lb_f32x16 b2;
    float w3[16];
    float b3;
};

__device__ inline void lb_load_wave_weights(const uint16_t *__restrict__ w, LbWaveWeights &W)
{
    const int lane = threadIdx.x & 63;
    const int row = lane & 31, half = lane >> 5;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i0 = lb_in_slot(c, 0, e), i1 = lb_in_slot(c, 1, e);
            const int in = half ? i1 : i0;
            W.a1[c][e] = __builtin_bit_cast(__bf16, in >= 0 ? w[LB_OFF_W1 + row * LB_NIN + max(in, 0)] : (uint16_t)0);
            // layer-2 column: the hidden unit of the lane's result 8c + e
            W.a2[c][e] = __builtin_bit_cast(__bf16, w[LB_OFF_W2 + row * LB_H + lb_row(8 * c + e, half)]);
        }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        W.b2[v] = bf2f(w[LB_OFF_B2 + lb_row(v, half)]);
        W.w3[v] = bf2f(w[LB_OFF_W3 + lb_row(v, half)]);
    }
    W.b3 = bf2f(w[LB_OFF_B3]);
}

// max(x, 0) and max(a, b) for finite values as one v_med3_f32 (fmaxf on a
// product's result adds a quieting v_max first).
__device__ inline float lb_relu(float x) { return __builtin_amdgcn_fmed3f(x, 0.f, 3.0e38f); }
__device__ inline float lb_fmax(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 3.0e38f); }

// Sum of a value over the lane's point in both half-waves (lane p and
// p + 32), the lower half's value first: the same float in every lane.
__device__ inline float lb_pair_sum(float x)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ inline float lb_pair_max(float x)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return lb_fmax(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

// y for the wave's 32 points.  F: this lane's 9 scaled field inputs (its
// half-wave's level), xv: the velocity feature, b1: the layer-1 biases of
// this lane's 16 result rows (the point's depth).  Returns y (every lane).
__device__ inline float lb_mlp_wave(const LbWaveWeights &W, const float F[9], float xv, const lb_f32x16 &b1)
{
    lb_bf16x8 bh, bl;
    lb_split8(F, bh, bl);
    lb_f32x16 acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a1[0], bh, b1, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a1[0], bl, acc, 0, 0, 0);
    const float t[8] = {F[8], xv, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    lb_split8(t, bh, bl);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a1[1], bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a1[1], bl, acc, 0, 0, 0);
    float h1[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) h1[v] = lb_relu(acc[v]);
    lb_split8(h1, bh, bl);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a2[0], bh, W.b2, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a2[0], bl, acc, 0, 0, 0);
    lb_split8(h1 + 8, bh, bl);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a2[1], bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a2[1], bl, acc, 0, 0, 0);
    float y = 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) y = fmaf(W.w3[v], lb_relu(acc[v]), y);
    return lb_pair_sum(y) + W.b3;
}

// The scaling of lb_finish_inputs on a lane's 9 field inputs, already
// offset by h and zero beyond the region's edge: s over both half-waves
// (both levels), inputs times 1/s (approximate reciprocal).  Returns s.
__device__ inline float lb_scale_lane_inputs(float F[9])
{
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) m = lb_fmax(m, fabsf(F[q]));
    const float s = lb_pair_max(m);
    const float inv = s > LB_S_TINY ? __builtin_amdgcn_rcpf(s) : 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) F[q] *= inv;
    return s;
}

// Layer-1 biases of this lane's result rows for depth d (LDS, LB_W x LB_H).
__device__ inline lb_f32x16 lb_bias1(const float *b1s, int d)
{
    const int half = (threadIdx.x & 63) >> 5;
    lb_f32x16 b;
#pragma unroll
    for (int v = 0; v < 16; ++v) b[v] = b1s[d * LB_H + lb_row(v, half)];
    return b;
}

// Correction pass, staged: a block takes 256 consecutive (u, v) points of
// one face (flat index u nv + v, so no tile is partly empty but the face's
// last) at the LB_W depths.  Thread (u, v) copies
// the LB_W + LB_R planes along the normal that lie inside the grid (of the
// 2 LB_R + LB_W that the tile's normal stencils reach) of both levels into
// LDS, offset by h (the planes beyond the grid's edge are zero: those
// inputs are), and vp c1 and its square root at the layer points.  LDS plane
// P is inward index P - R from the face (P < R: beyond the edge), so input m
// of depth d is plane d + m on either side.  Wavefront w then evaluates
// depth w: LB_TU groups of 32 consecutive points.
constexpr int LB_TU = 8, LB_TV = 32;
constexpr int LB_NPL = 2 * LB_R + LB_W;  // normal planes of a tile (12)
constexpr int LB_NIG = LB_W + LB_R;      // of which inside the grid (8)
static_assert(LB_TV == 32 && LB_W == 4 && LB_TU * LB_TV == 256, "one thread per (u, v); one wavefront per depth");

struct LbCorrectTileArgs {
    const float *pn, *pnm1, *vp;
    FdLayout L;
    const uint16_t *w;
    float *corr;
    int64_t off[LB_FACES + 1];
    int blk0[LB_FACES + 1];  // first block of each face
    float h, c1;
    float ymax[LB_W];        // bound of y at each depth (lb.hpp)
    const float *kvcs;       // k-faces: vp c1 and sqrt(vp c1), compact (or null: from vp)
    const float *kslab_in;   // k-faces: previous level's planes, compact (or null: from pnm1)
    float *kslab_out;        // k-faces: the current level's planes written here (or null)
    int npts_k;
};

__global__ void __launch_bounds__(256) lb_correct_tile_kernel(LbCorrectTileArgs A)
{
    __shared__ float b1s[LB_W * LB_H];
    __shared__ float S[2][LB_NPL][LB_TU][LB_TV];
    __shared__ float Svc[LB_W][LB_TU][LB_TV];  // vp c1 at the tile's layer points [depth][u][v]
    __shared__ float Svs[LB_W][LB_TU][LB_TV];  // sqrt(vp c1), the velocity feature
    int f = 0;
    while (f + 1 < LB_FACES && (int)blockIdx.x >= A.blk0[f + 1]) ++f;
    const int n[3] = {A.L.nx, A.L.ny, A.L.nz};
    const LbFaceGeom g = lb_face(n, f);
    const int a = g.axis, ua = (a == 0) ? 1 : 0, va = (a == 2) ? 1 : 2;
    const int npts = g.nu * g.nv;
    const int flat0 = ((int)blockIdx.x - A.blk0[f]) * (LB_TU * LB_TV);
    const uint16_t *w = A.w + f * LB_NP;
    static_assert(LB_W * LB_H <= 256, "one layer-1 bias per thread");
    const uint16_t b1raw = w[LB_OFF_B1 + min((int)threadIdx.x, LB_W * LB_H - 1)];
    // in-grid plane q (normal index q, or n - W - R + q at the high face)
    // is LDS plane R + q (low face) or W + 2R - 1 - q (high face)
    const int tu = (int)threadIdx.x / LB_TV, tv = (int)threadIdx.x % LB_TV;
    const int flat = min(flat0 + (int)threadIdx.x, npts - 1);
    int c[3];
    c[a] = g.side ? n[a] - LB_NIG : 0;
    c[ua] = flat / g.nv;
    c[va] = flat - c[ua] * g.nv;
    const int64_t o0 = fd_off(A.L, c[0], c[1], c[2]);
    const int64_t step = a == 0 ? A.L.pitch_j : (a == 1 ? A.L.pitch_k : 1);
    // every load is issued before the first is waited for.  On the k-faces
    // (a row end costs a whole line) the previous level and the velocity
    // terms come from compact stores when given, and the current level's
    // planes are written to one for the next pass.
    const bool kf = a == 2;
    const int64_t kb = (int64_t)(f - 4) * LB_NIG * A.npts_k + flat;  // k-face store element (q = 0)
    const float *kin = kf ? A.kslab_in : nullptr;
    float *kout = kf ? A.kslab_out : nullptr;
    const bool kvc = kf && A.kvcs;
    float r0[LB_NIG], r1[LB_NIG], rv[LB_W], rs[LB_W];
#pragma unroll
    for (int q = 0; q < LB_NIG; ++q) {
        r0[q] = A.pn[o0 + q * step];
        r1[q] = kin ? kin[kb + (int64_t)q * A.npts_k] : A.pnm1[o0 + q * step];
    }
    if (kvc) {
        const int64_t vb = (int64_t)(f - 4) * 2 * LB_W * A.npts_k + flat;
#pragma unroll
        for (int dd = 0; dd < LB_W; ++dd) {
            rv[dd] = A.kvcs[vb + (int64_t)dd * A.npts_k];
            rs[dd] = A.kvcs[vb + (int64_t)(LB_W + dd) * A.npts_k];
        }
    } else {
#pragma unroll
        for (int dd = 0; dd < LB_W; ++dd) rv[dd] = A.vp[o0 + (g.side ? LB_NIG - 1 - dd : dd) * step];
    }
    LbWaveWeights W;  // its loads in flight with the staging's
    lb_load_wave_weights(w, W);
    if (kout && flat0 + (int)threadIdx.x < npts) {
#pragma unroll
        for (int q = 0; q < LB_NIG; ++q) kout[kb + (int64_t)q * A.npts_k] = r0[q];
    }
#pragma unroll
    for (int q = 0; q < LB_NIG; ++q) {
        const int P = g.side ? LB_NPL - 1 - q : LB_R + q;
        S[0][P][tu][tv] = r0[q] - A.h;
        S[1][P][tu][tv] = r1[q] - A.h;
    }
#pragma unroll
    for (int P = 0; P < LB_R; ++P) {  // beyond the edge
        S[0][P][tu][tv] = 0.f;
        S[1][P][tu][tv] = 0.f;
    }
#pragma unroll
    for (int dd = 0; dd < LB_W; ++dd) {
        const float vc = kvc ? rv[dd] : rv[dd] * A.c1;
        Svc[dd][tu][tv] = vc;
        Svs[dd][tu][tv] = kvc ? rs[dd] : sqrtf(vc);
    }
    if (threadIdx.x < LB_W * LB_H) b1s[threadIdx.x] = bf2f(b1raw);
    __syncthreads();
    const int d = threadIdx.x >> 6;  // this wavefront's depth
    const int p = threadIdx.x & 31;  // this lane's point in the group (both half-waves)
    const int half = (threadIdx.x >> 5) & 1;
    const float ymax = A.ymax[d];
    float *corr = A.corr + A.off[f] + (int64_t)d * npts;
#pragma unroll 1
    for (int uu = 0; uu < LB_TU; ++uu) {
        const int fg = flat0 + uu * LB_TV;  // the group's first point
        if (fg >= npts) break;  // uniform per wavefront
        float F[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) F[m] = S[half][d + m][uu][p];  // offset m - 4 along the inward normal
        const float s = lb_scale_lane_inputs(F);
        const float y = __builtin_amdgcn_fmed3f(lb_mlp_wave(W, F, Svs[d][uu][p], lb_bias1(b1s, d)), -ymax, ymax);
        if (half == 0 && fg + p < npts) corr[fg + p] = Svc[d][uu][p] * s * y;
    }
}

// MFMA inference on given inputs (testing): y[n] for samples of one face,
// inputs as lb_inputs forms them.
__global__ void __launch_bounds__(256) lb_infer_kernel(const uint16_t *w, const float *samples, int64_t n, float *y)
{
    __shared__ float b1s[LB_W * LB_H];
    for (int t = threadIdx.x; t < LB_W * LB_H; t += blockDim.x) b1s[t] = bf2f(w[LB_OFF_B1 + t]);
    __syncthreads();
    const int64_t wv = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv * 32 >= n) return;
    LbWaveWeights W;
    lb_load_wave_weights(w, W);
    const int64_t idx = wv * 32 + (threadIdx.x & 31);
    const int64_t ci = idx < n ? idx : n - 1;
    const int half = (threadIdx.x >> 5) & 1;
    float F[9];
    for (int m = 0; m < 9; ++m) F[m] = samples[ci * LB_NSTRIDE + 9 * half + m];
    const float xv = samples[ci * LB_NSTRIDE + LB_NIN - 1];
    const int d = min(max((int)samples[ci * LB_NSTRIDE + LB_NIN + 1], 0), LB_W - 1);
    const float r = lb_mlp_wave(W, F, xv, lb_bias1(b1s, d));
    if (half == 0 && idx < n) y[idx] = r;
}

// Apply pass: one thread per point of the union of the layers; adds the
// corrections of every face whose layer holds the point.  One launch walks
// the union in three parts, each with its fastest index along k: the
// i-layers (2 LB_W planes of ny x nz points), the j-layers with i interior
// (rows along k), and the k-layers with i and j interior (the 2 LB_W points
// of each (i, j) column, both ends).
struct LbApplyArgs {
    float *p;
    FdLayout L;
    const float *corr;
    int64_t off[LB_FACES + 1];
    int nA, nB, nC;     // points of the three parts
    float ldamp[LB_W];  // damping factor of the new value per layer depth (1: none)
};

__global__ void __launch_bounds__(256) lb_apply_kernel(LbApplyArgs A)
{
    const int nx = A.L.nx, ny = A.L.ny, nz = A.L.nz;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < A.nA) {
        k = t % nz;
        const int r = t / nz, q = r / ny;
        j = r - q * ny;
        i = q < LB_W ? q : nx - 2 * LB_W + q;
    } else if ((t -= A.nA) < A.nB) {
        k = t % nz;
        const int r = t / nz, q = r % (2 * LB_W);
        i = LB_W + r / (2 * LB_W);
        j = q < LB_W ? q : ny - 2 * LB_W + q;
    } else if ((t -= A.nB) < A.nC) {
        const int q = t % (2 * LB_W), r = t / (2 * LB_W), nyi = ny - 2 * LB_W;
        j = LB_W + r % nyi;
        i = LB_W + r / nyi;
        k = q < LB_W ? q : nz - 2 * LB_W + q;
    } else {
        return;
    }
    const int n[3] = {nx, ny, nz};
    const int c[3] = {i, j, k};
    float add = 0.f, damp = 1.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool lo = c[a] < LB_W, hi = c[a] >= n[a] - LB_W;
        const int side = hi ? 1 : 0;
        const int d = lo ? c[a] : (hi ? n[a] - 1 - c[a] : 0);  // 0 when not in this axis's layers
        const int nu = (a == 0) ? ny : nx;
        const int nv = (a == 2) ? ny : nz;
        const int u = (a == 0) ? j : i;
        const int v = (a == 2) ? j : k;
        if (lo || hi) {  // only the faces whose layer holds the point are read
            add += A.corr[A.off[2 * a + side] + ((int64_t)d * nu + u) * nv + v];
            damp *= A.ldamp[d];
        }
    }
    const int64_t o = fd_off(A.L, i, j, k);
    A.p[o] = mul_nc(add_nc(A.p[o], add), damp);
}

// vp c1 and sqrt(vp c1) at the k-faces' layer points (the same operations as
// the correction pass's staging): kvcs[((f - 4) * 2 + which) * LB_W + d][i ny + j].
__global__ void lb_kface_vp_kernel(const float *vp, FdLayout L, float c1, float *kvcs)
{
    const int64_t npts = (int64_t)L.nx * L.ny;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * LB_W * npts) return;
    const int fk = (int)(t / (LB_W * npts)), d = (int)((t / npts) % LB_W);
    const int64_t flat = t % npts;
    const int i = (int)(flat / L.ny), j = (int)(flat % L.ny);
    const int k = fk ? L.nz - 1 - d : d;
    const float vc = vp[fd_off(L, i, j, k)] * c1;
    kvcs[((int64_t)fk * 2 * LB_W + d) * npts + flat] = vc;
    kvcs[((int64_t)fk * 2 * LB_W + LB_W + d) * npts + flat] = sqrtf(vc);
}

hipError_t lb_enqueue_kface_vp(const LbInfer &lb, const FdLayout &L, const float *vp, float *kvcs,
                               hipStream_t stream)
{
    const int64_t n = 2 * LB_W * (int64_t)L.nx * L.ny;
    hipLaunchKernelGGL(lb_kface_vp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, vp, L, lb.c1,
                       kvcs);
    return hipGetLastError();
}

hipError_t lb_enqueue_correct(const LbInfer &lb, const FdLayout &L, const float *pn, const float *pnm1,
                              const float *vp, hipStream_t stream, const float *kslab_in, float *kslab_out)
{
    // grids under the learned layer have at least 4 LB_W points per side
    // (phc_fd_set_learned_boundary), so every tile's box of normal planes
    // lies inside the grid and its halo
    LbCorrectTileArgs T;
    T.pn = pn;
    T.pnm1 = pnm1;
    T.vp = vp;
    T.L = L;
    T.w = lb.w;
    T.corr = lb.corr;
    T.h = lb.halo;
    T.c1 = lb.c1;
    T.blk0[0] = 0;
    T.kvcs = lb.kvcs;
    T.kslab_in = kslab_in;
    T.kslab_out = kslab_out;
    T.npts_k = L.nx * L.ny;
    for (int d = 0; d < LB_W; ++d) T.ymax[d] = lb_ymax(lb.chat, d);
    for (int f = 0; f <= LB_FACES; ++f) T.off[f] = lb.corr_off[f];
    const int nn[3] = {L.nx, L.ny, L.nz};
    for (int f = 0; f < LB_FACES; ++f) {
        const LbFaceGeom g = lb_face(nn, f);
        T.blk0[f + 1] = T.blk0[f] + (g.nu * g.nv + LB_TU * LB_TV - 1) / (LB_TU * LB_TV);
    }
    hipLaunchKernelGGL(lb_correct_tile_kernel, dim3((unsigned)T.blk0[LB_FACES]), dim3(256), 0, stream, T);
    return hipGetLastError();
}

hipError_t lb_enqueue_apply(const LbInfer &lb, const FdLayout &L, float *pnew, hipStream_t stream)
{
    LbApplyArgs A;
    A.p = pnew;
    A.L = L;
    A.corr = lb.corr;
    for (int f = 0; f <= LB_FACES; ++f) A.off[f] = lb.corr_off[f];
    for (int d = 0; d < LB_W; ++d) A.ldamp[d] = lb.ldamp[d];
    // (every side is at least 4 LB_W points, phc_fd_set_learned_boundary)
    A.nA = 2 * LB_W * L.ny * L.nz;
    A.nB = (L.nx - 2 * LB_W) * 2 * LB_W * L.nz;
    A.nC = (L.nx - 2 * LB_W) * (L.ny - 2 * LB_W) * 2 * LB_W;
    const int64_t tot = (int64_t)A.nA + A.nB + A.nC;
    if (tot == 0) return hipSuccess;
    hipLaunchKernelGGL(lb_apply_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, A);
    return hipGetLastError();
}

} // namespace phc

using namespace phc;

struct phc_lb {
    int device = 0;  // the device current at create (the stream's device)
    DevBuf theta, grad, m, v, loss, wbf, idx;
    DevBuf data[LB_FACES];
    DevBuf count;  // unsigned long long per face
    int64_t cap = 0;
    int64_t adam_t = 0;
    float noise = 0.f;  // training input noise (phc_lb_set_train_noise)
    hipStream_t stream = nullptr;
};

// ---------------------------------------------------- training without allocation
// phc_lb_train_noalloc keeps phc_lb_train's per-call buffers (minibatch
// indices, per-block partial gradients and losses, loss history) in one static
// device array, `scratch`, of fixed size: LB_NA_MAX_BATCH samples per face and
// LB_NA_MAX_STEPS steps per call, 5.6 MiB.  Each GPU holds its own copy until
// the process ends; its address on a device is resolved once, at the first
// phc_lb_create with that device current, and kept in lb_na_ptr.  Calls on one
// GPU share the array: calls that overlap on one GPU race.
constexpr int LB_NA_MAX_BATCH = 8192;
constexpr int LB_NA_MAX_STEPS = 16384;
constexpr int LB_NA_MAX_BLOCKS = LB_FACES * ((LB_NA_MAX_BATCH + LB_TB - 1) / LB_TB);
constexpr size_t lb_na_round(size_t n) { return (n + 63) / 64 * 64; }  // 256-byte regions
constexpr size_t LB_NA_PART = 0;
constexpr size_t LB_NA_PLOSS = LB_NA_PART + lb_na_round((size_t)LB_NA_MAX_BLOCKS * LB_NP);
constexpr size_t LB_NA_LOSS = LB_NA_PLOSS + lb_na_round(LB_NA_MAX_BLOCKS);
constexpr size_t LB_NA_IDX = LB_NA_LOSS + lb_na_round(LB_NA_MAX_STEPS);
constexpr size_t LB_NA_FLOATS = LB_NA_IDX + lb_na_round((size_t)LB_FACES * LB_NA_MAX_BATCH);
static_assert(sizeof(int) == sizeof(float), "indices share the float array");

static __device__ __align__(256) float scratch[LB_NA_FLOATS];

constexpr int LB_NA_MAX_DEVICES = 64;
static std::mutex lb_na_mutex;
static float *lb_na_ptr[LB_NA_MAX_DEVICES];

// The device address of `scratch` on device dev, which is current.
static phc_status lb_na_setup(int dev)
{
    if (dev < 0 || dev >= LB_NA_MAX_DEVICES) return fail(PHC_ERR_INVALID, "device %d outside 0..%d", dev,
                                                         LB_NA_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> g(lb_na_mutex);
    if (lb_na_ptr[dev]) return PHC_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(PHC_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    lb_na_ptr[dev] = static_cast<float *>(p);
    return PHC_OK;
}

static float *lb_na_base(int dev)
{
    if (dev < 0 || dev >= LB_NA_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(lb_na_mutex);
    return lb_na_ptr[dev];
}

static phc_status lb_upload_bf16(phc_lb *h)
{
    const int n = LB_FACES * LB_NP;
    hipLaunchKernelGGL(lb_to_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->theta.as<float>(),
                       h->wbf.as<uint16_t>(), n);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

static phc_status lb_counts(phc_lb *h, int64_t c[LB_FACES])
{
    unsigned long long hc[LB_FACES];
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemcpy(hc, h->count.ptr, sizeof(hc), hipMemcpyDeviceToHost));
    for (int f = 0; f < LB_FACES; ++f) c[f] = std::min<int64_t>((int64_t)hc[f], h->cap);
    return PHC_OK;
}

// one forward/backward pass of a batch; loss is accumulated into h->loss
static phc_status lb_launch_train(phc_lb *h, const LbBatch &B)
{
    if (B.block_start[LB_FACES] == 0) return PHC_OK;
    hipLaunchKernelGGL(lb_train_kernel, dim3(B.block_start[LB_FACES]), dim3(LB_TB), 0, h->stream,
                       h->theta.as<float>(), h->grad.as<float>(), h->loss.as<float>(), B);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

static void lb_fill(phc_lb *h, float *p, int n, float val)
{
    hipLaunchKernelGGL(lb_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, h->stream, p, n, val);
}

extern "C" {

phc_status phc_lb_create(uint64_t seed, phc_lb **out)
{
    if (!out) return fail(PHC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    phc_lb *h = new phc_lb();
    auto bail = [&](phc_status s) { phc_lb_destroy(h); return s; };
    if (hipGetDevice(&h->device) != hipSuccess) return bail(fail(PHC_ERR_HIP, "hipGetDevice failed"));
    phc_status sst = lb_na_setup(h->device);
    if (sst != PHC_OK) return bail(sst);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "hipStreamCreate failed"));
    const size_t n = LB_FACES * LB_NP;
    phc_status st;
    if ((st = h->theta.alloc(4 * n)) || (st = h->grad.alloc(4 * n)) || (st = h->m.alloc(4 * n)) ||
        (st = h->v.alloc(4 * n)) || (st = h->loss.alloc(4)) || (st = h->wbf.alloc(2 * n)) ||
        (st = h->count.alloc(8 * LB_FACES)))
        return bail(st);
    // He-uniform initial weights, zero biases, small output layer
    std::mt19937_64 rng(seed);
    std::vector<float> th(n, 0.f);
    for (int f = 0; f < LB_FACES; ++f) {
        float *p = th.data() + f * LB_NP;
        std::uniform_real_distribution<float> u1(-std::sqrt(6.f / LB_NIN), std::sqrt(6.f / LB_NIN));
        std::uniform_real_distribution<float> u2(-std::sqrt(6.f / LB_H), std::sqrt(6.f / LB_H));
        std::uniform_real_distribution<float> u3(-0.1f / std::sqrt((float)LB_H), 0.1f / std::sqrt((float)LB_H));
        for (int q = 0; q < LB_H * LB_NIN; ++q) p[LB_OFF_W1 + q] = u1(rng);
        for (int q = 0; q < LB_H * LB_H; ++q) p[LB_OFF_W2 + q] = u2(rng);
        for (int q = 0; q < LB_H; ++q) p[LB_OFF_W3 + q] = u3(rng);
    }
    if (hipMemcpy(h->theta.ptr, th.data(), 4 * n, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "upload failed"));
    std::vector<float> z(n, 0.f);
    if (hipMemcpy(h->grad.ptr, z.data(), 4 * n, hipMemcpyHostToDevice) ||
        hipMemcpy(h->m.ptr, z.data(), 4 * n, hipMemcpyHostToDevice) ||
        hipMemcpy(h->v.ptr, z.data(), 4 * n, hipMemcpyHostToDevice) ||
        hipMemset(h->count.ptr, 0, 8 * LB_FACES) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "upload failed"));
    if ((st = lb_upload_bf16(h))) return bail(st);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fail(PHC_ERR_HIP, "sync failed"));
    *out = h;
    return PHC_OK;
}

phc_status phc_lb_destroy(phc_lb *h)
{
    if (!h) return PHC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (DevBuf *b : {&h->theta, &h->grad, &h->m, &h->v, &h->loss, &h->wbf, &h->idx, &h->count}) b->release();
    for (auto &b : h->data) b.release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PHC_OK;
}

phc_status phc_lb_get_params(phc_lb *h, float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "NULL argument");
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemcpy(host, h->theta.ptr, 4 * LB_FACES * LB_NP, hipMemcpyDeviceToHost));
    return PHC_OK;
}

phc_status phc_lb_set_params(phc_lb *h, const float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "NULL argument");
    for (int q = 0; q < LB_FACES * LB_NP; ++q)
        if (!std::isfinite(host[q])) return fail(PHC_ERR_INVALID, "parameter %d is not finite", q);
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemcpy(h->theta.ptr, host, 4 * LB_FACES * LB_NP, hipMemcpyHostToDevice));
    PHC_TRY(lb_upload_bf16(h));
    h->adam_t = 0;
    std::vector<float> z(LB_FACES * LB_NP, 0.f);
    PHC_HIP(hipMemcpy(h->m.ptr, z.data(), 4 * z.size(), hipMemcpyHostToDevice));
    PHC_HIP(hipMemcpy(h->v.ptr, z.data(), 4 * z.size(), hipMemcpyHostToDevice));
    PHC_HIP(hipMemcpy(h->grad.ptr, z.data(), 4 * z.size(), hipMemcpyHostToDevice));
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

phc_status phc_lb_set_capacity(phc_lb *h, int64_t per_face)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (per_face < 1 || per_face > ((int64_t)1 << 31) - 1)
        return fail(PHC_ERR_INVALID, "capacity %lld outside 1..2^31-1", (long long)per_face);
    PHC_HIP(hipStreamSynchronize(h->stream));
    for (auto &b : h->data) PHC_TRY(b.alloc(sizeof(float) * LB_NSTRIDE * (size_t)per_face));
    h->cap = per_face;
    PHC_HIP(hipMemset(h->count.ptr, 0, 8 * LB_FACES));
    return PHC_OK;
}

phc_status phc_lb_clear_samples(phc_lb *h)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemset(h->count.ptr, 0, 8 * LB_FACES));
    return PHC_OK;
}

phc_status phc_lb_sample_counts(phc_lb *h, int64_t *counts)
{
    if (!h || !counts) return fail(PHC_ERR_INVALID, "NULL argument");
    return lb_counts(h, counts);
}

phc_status phc_lb_get_samples(phc_lb *h, int face, int64_t first, int64_t n, float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "NULL argument");
    if (face < 0 || face >= LB_FACES) return fail(PHC_ERR_INVALID, "face %d outside 0..5", face);
    int64_t c[LB_FACES];
    PHC_TRY(lb_counts(h, c));
    if (first < 0 || n < 0 || first + n > c[face])
        return fail(PHC_ERR_INVALID, "samples %lld..%lld outside the %lld of face %d", (long long)first,
                    (long long)(first + n), (long long)c[face], face);
    if (n > 0)
        PHC_HIP(hipMemcpy(host, h->data[face].as<float>() + first * LB_NSTRIDE, 4 * LB_NSTRIDE * n,
                          hipMemcpyDeviceToHost));
    return PHC_OK;
}

phc_status phc_lb_add_samples(phc_lb *h, int face, const float *host, int64_t n)
{
    if (!h || (!host && n > 0)) return fail(PHC_ERR_INVALID, "NULL argument");
    if (face < 0 || face >= LB_FACES) return fail(PHC_ERR_INVALID, "face %d outside 0..5", face);
    int64_t c[LB_FACES];
    PHC_TRY(lb_counts(h, c));
    if (n < 0 || c[face] + n > h->cap)
        return fail(PHC_ERR_INVALID, "%lld samples exceed the capacity %lld", (long long)(c[face] + n),
                    (long long)h->cap);
    if (n > 0)
        PHC_HIP(hipMemcpy(h->data[face].as<float>() + c[face] * LB_NSTRIDE, host, 4 * LB_NSTRIDE * n,
                          hipMemcpyHostToDevice));
    const unsigned long long nc = (unsigned long long)(c[face] + n);
    PHC_HIP(hipMemcpy(h->count.as<unsigned long long>() + face, &nc, 8, hipMemcpyHostToDevice));
    return PHC_OK;
}

phc_status phc_lb_collect(phc_lb *h, phc_fd *wide, int cur, const int *origin, const int *dims, float halo,
                          float keep, float s_min, uint64_t seed)
{
    RoctxRange rr_("phc_lb_collect");
    if (!h || !wide || !origin || !dims) return fail(PHC_ERR_INVALID, "NULL argument");
    if (cur < 0 || cur > 1) return fail(PHC_ERR_INVALID, "cur level %d is not 0 or 1", cur);
    if (h->cap < 1) return fail(PHC_ERR_STATE, "no sample store: call phc_lb_set_capacity first");
    FdFieldView fv;
    PHC_TRY(fd_field_view(wide, &fv));
    const int N[3] = {fv.L.nx, fv.L.ny, fv.L.nz};
    for (int a = 0; a < 3; ++a) {
        if (dims[a] < 2 * LB_W + 1) return fail(PHC_ERR_INVALID, "region side %d too small", dims[a]);
        // the region, its outward stencil reach and the update's reach lie outside the sponge
        if (origin[a] - LB_R < fv.sw || origin[a] + dims[a] + LB_R > N[a] - fv.sw)
            return fail(PHC_ERR_INVALID, "region plus %d points must stay outside the %d-point sponge", LB_R,
                        fv.sw);
    }
    LbCollectArgs A;
    A.G.pn = fv.lev[cur];
    A.G.pnm1 = fv.lev[1 - cur];
    A.G.vp = fv.vp;
    A.G.L = fv.L;
    for (int a = 0; a < 3; ++a) {
        A.G.o[a] = origin[a];
        A.G.n[a] = dims[a];
    }
    A.h = halo;
    A.c1 = fv.c1;
    for (int l = 0; l <= LB_R; ++l) A.chat[l] = fv.chat[l];
    A.keep = keep;
    A.s_min = s_min;
    A.seed = seed;
    for (int f = 0; f < LB_FACES; ++f) A.out[f] = h->data[f].as<float>();
    A.count = h->count.as<unsigned long long>();
    A.cap = h->cap;
    A.face_off[0] = 0;
    for (int f = 0; f < LB_FACES; ++f) {
        const LbFaceGeom g = lb_face(A.G.n, f);
        A.face_off[f + 1] = A.face_off[f] + (int64_t)LB_W * g.nu * g.nv;
    }
    PHC_HIP(hipStreamSynchronize(fv.stream));
    hipLaunchKernelGGL(lb_collect_kernel, dim3(LB_FACES), dim3(LB_CB), 0, h->stream, A);
    PHC_HIP(hipGetLastError());
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

phc_status phc_lb_collect_onpolicy(phc_lb *h, phc_fd *thin, int thin_cur, phc_fd *wide, int wide_cur,
                                   const int *origin, float halo, float keep, float s_min, uint64_t seed)
{
    RoctxRange rr_("phc_lb_collect_onpolicy");
    if (!h || !thin || !wide || !origin) return fail(PHC_ERR_INVALID, "NULL argument");
    if (thin_cur < 0 || thin_cur > 1 || wide_cur < 0 || wide_cur > 1)
        return fail(PHC_ERR_INVALID, "levels must be 0 or 1");
    if (h->cap < 1) return fail(PHC_ERR_STATE, "no sample store: call phc_lb_set_capacity first");
    FdFieldView tv, wv;
    PHC_TRY(fd_field_view(thin, &tv));
    PHC_TRY(fd_field_view(wide, &wv));
    if (tv.c1 != wv.c1) return fail(PHC_ERR_INVALID, "thin and wide runs use different stencils");
    const int n[3] = {tv.L.nx, tv.L.ny, tv.L.nz}, N[3] = {wv.L.nx, wv.L.ny, wv.L.nz};
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 4 * LB_W) return fail(PHC_ERR_INVALID, "thin grid side %d below %d", n[a], 4 * LB_W);
        if (origin[a] - LB_R < wv.sw || origin[a] + n[a] + LB_R > N[a] - wv.sw)
            return fail(PHC_ERR_INVALID, "the thin region plus %d points must lie outside the wide run's sponge",
                        LB_R);
    }
    LbOnPolicyArgs A;
    A.T.pn = tv.lev[thin_cur];
    A.T.pnm1 = tv.lev[1 - thin_cur];
    A.T.vp = tv.vp;
    A.T.L = tv.L;
    for (int a = 0; a < 3; ++a) {
        A.T.o[a] = 0;
        A.T.n[a] = n[a];
        A.wo[a] = origin[a];
    }
    A.wcur = wv.lev[wide_cur];
    A.LW = wv.L;
    A.h = halo;
    A.c1 = tv.c1;
    for (int l = 0; l <= LB_R; ++l) A.chat[l] = tv.chat[l];
    A.keep = keep;
    A.s_min = s_min;
    A.t_max = 16.f;
    A.seed = seed;
    for (int f = 0; f < LB_FACES; ++f) A.out[f] = h->data[f].as<float>();
    A.count = h->count.as<unsigned long long>();
    A.cap = h->cap;
    A.face_off[0] = 0;
    for (int f = 0; f < LB_FACES; ++f) {
        const LbFaceGeom g = lb_face(n, f);
        A.face_off[f + 1] = A.face_off[f] + (int64_t)LB_W * g.nu * g.nv;
    }
    PHC_HIP(hipStreamSynchronize(tv.stream));
    PHC_HIP(hipStreamSynchronize(wv.stream));
    hipLaunchKernelGGL(lb_onpolicy_kernel, dim3(LB_FACES), dim3(LB_CB), 0, h->stream, A);
    PHC_HIP(hipGetLastError());
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

// Loss (mean squared error of y against the targets over every stored
// sample) and, if grad is not NULL, its gradient with respect to the FP32
// parameters (evaluated at their bf16 values).
phc_status phc_lb_eval(phc_lb *h, float *loss, float *grad)
{
    if (!h || !loss) return fail(PHC_ERR_INVALID, "NULL argument");
    int64_t c[LB_FACES];
    PHC_TRY(lb_counts(h, c));
    int64_t tot = 0;
    for (int f = 0; f < LB_FACES; ++f) tot += c[f];
    if (tot == 0) return fail(PHC_ERR_STATE, "no samples");
    LbBatch B;
    B.noise = 0.f;
    B.nseed = 0;
    B.partial = B.ploss = nullptr;
    B.block_start[0] = 0;
    for (int f = 0; f < LB_FACES; ++f) {
        B.samples[f] = h->data[f].as<float>();
        B.idx[f] = nullptr;
        B.count[f] = (int)c[f];
        B.block_start[f + 1] = B.block_start[f] + (int)((c[f] + LB_TB - 1) / LB_TB);
    }
    B.inv_total = 1.f / (float)tot;
    const int np = LB_FACES * LB_NP;
    DevBuf saved;
    PHC_TRY(saved.alloc(4 * (size_t)np));  // keep the gradient a training run may be accumulating
    PHC_HIP(hipMemcpyAsync(saved.ptr, h->grad.ptr, 4 * np, hipMemcpyDeviceToDevice, h->stream));
    lb_fill(h, h->grad.as<float>(), np, 0.f);
    lb_fill(h, h->loss.as<float>(), 1, 0.f);
    PHC_TRY(lb_launch_train(h, B));
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemcpy(loss, h->loss.ptr, 4, hipMemcpyDeviceToHost));
    if (grad) PHC_HIP(hipMemcpy(grad, h->grad.ptr, 4 * np, hipMemcpyDeviceToHost));
    PHC_HIP(hipMemcpy(h->grad.ptr, saved.ptr, 4 * np, hipMemcpyDeviceToDevice));
    return PHC_OK;
}

phc_status phc_lb_set_train_noise(phc_lb *h, float std)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (!(std >= 0.f) || std > 1.f) return fail(PHC_ERR_INVALID, "noise std %g outside [0, 1]", std);
    h->noise = std;
    return PHC_OK;
}

// The training loop of phc_lb_train on given per-call buffers: idx (LB_FACES x
// batch_per_face ints), lh (steps floats), part (per-block partial gradients,
// LB_NP floats per block), pl (per-block losses).  Every buffer is written
// before it is read, so none needs clearing.
static phc_status lb_train_core(phc_lb *h, const int64_t *c, int steps, int batch_per_face, float lr, uint64_t seed,
                                int *idx, float *lh, float *part, float *pl, float *loss_history)
{
    int nf = 0;
    for (int f = 0; f < LB_FACES; ++f) nf += c[f] > 0;
    LbBatch B;
    B.block_start[0] = 0;
    for (int f = 0; f < LB_FACES; ++f) {
        B.samples[f] = h->data[f].as<float>();
        B.idx[f] = idx + (size_t)f * batch_per_face;
        B.count[f] = c[f] > 0 ? batch_per_face : 0;
        B.block_start[f + 1] = B.block_start[f] + (B.count[f] + LB_TB - 1) / LB_TB;
    }
    B.inv_total = 1.f / ((float)batch_per_face * nf);
    const int np = LB_FACES * LB_NP;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    B.partial = part;
    B.ploss = pl;
    for (int s = 0; s < steps; ++s) {
        for (int f = 0; f < LB_FACES; ++f)
            if (c[f] > 0)
                hipLaunchKernelGGL(lb_random_idx_kernel, dim3((batch_per_face + 255) / 256), dim3(256), 0, h->stream,
                                   idx + (size_t)f * batch_per_face, batch_per_face, (int)c[f],
                                   seed * 1000003ull + (uint64_t)s * LB_FACES + f);
        B.noise = h->noise;
        B.nseed = seed * 7919ull + (uint64_t)s;
        hipLaunchKernelGGL(lb_train_kernel, dim3(B.block_start[LB_FACES]), dim3(LB_TB), 0, h->stream,
                           h->theta.as<float>(), h->grad.as<float>(), lh + s, B);
        hipLaunchKernelGGL(lb_reduce_kernel, dim3((np + 255) / 256), dim3(256), 0, h->stream, B,
                           h->grad.as<float>(), lh + s);
        ++h->adam_t;
        const float c1 = 1.f / (1.f - std::pow(b1, (float)h->adam_t));
        const float c2 = 1.f / (1.f - std::pow(b2, (float)h->adam_t));
        hipLaunchKernelGGL(lb_adam_kernel, dim3((np + 255) / 256), dim3(256), 0, h->stream, h->theta.as<float>(),
                           h->grad.as<float>(), h->m.as<float>(), h->v.as<float>(), np, lr, b1, b2, c1, c2, eps);
    }
    PHC_HIP(hipGetLastError());
    PHC_TRY(lb_upload_bf16(h));
    PHC_HIP(hipStreamSynchronize(h->stream));
    if (loss_history && steps > 0)
        PHC_HIP(hipMemcpy(loss_history, lh, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return PHC_OK;
}

static phc_status lb_train_check(phc_lb *h, int steps, int batch_per_face, float lr, int64_t *c)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (steps < 0 || batch_per_face < 1 || batch_per_face > (1 << 24) || !(lr > 0.f))
        return fail(PHC_ERR_INVALID, "bad training parameters");
    PHC_TRY(lb_counts(h, c));
    int nf = 0;
    for (int f = 0; f < LB_FACES; ++f) nf += c[f] > 0;
    if (nf == 0) return fail(PHC_ERR_STATE, "no samples");
    return PHC_OK;
}

// Adam on minibatches of batch_per_face random samples of each face;
// loss_history[s] (may be NULL) = minibatch loss before step s.  The gradient
// of each step is summed per block and the blocks in a fixed order, so a run
// repeats bit for bit.
phc_status phc_lb_train(phc_lb *h, int steps, int batch_per_face, float lr, uint64_t seed, float *loss_history)
{
    RoctxRange rr_("phc_lb_train");
    int64_t c[LB_FACES];
    PHC_TRY(lb_train_check(h, steps, batch_per_face, lr, c));
    int nblocks = 0;
    for (int f = 0; f < LB_FACES; ++f) nblocks += c[f] > 0 ? (batch_per_face + LB_TB - 1) / LB_TB : 0;
    PHC_TRY(h->idx.alloc(sizeof(int) * (size_t)batch_per_face * LB_FACES));
    DevBuf lh, part, pl;
    PHC_TRY(lh.alloc(sizeof(float) * std::max(steps, 1)));
    PHC_TRY(part.alloc(sizeof(float) * (size_t)nblocks * LB_NP));
    PHC_TRY(pl.alloc(sizeof(float) * (size_t)nblocks));
    return lb_train_core(h, c, steps, batch_per_face, lr, seed, h->idx.as<int>(), lh.as<float>(), part.as<float>(),
                         pl.as<float>(), loss_history);
}

phc_status phc_lb_train_noalloc(phc_lb *h, int steps, int batch_per_face, float lr, uint64_t seed,
                                float *loss_history)
{
    int64_t c[LB_FACES];
    PHC_TRY(lb_train_check(h, steps, batch_per_face, lr, c));
    if (batch_per_face > LB_NA_MAX_BATCH)
        return fail(PHC_ERR_INVALID, "batch of %d samples per face exceeds the no-alloc limit %d", batch_per_face,
                    LB_NA_MAX_BATCH);
    if (steps > LB_NA_MAX_STEPS)
        return fail(PHC_ERR_INVALID, "%d steps exceed the no-alloc limit %d per call", steps, LB_NA_MAX_STEPS);
    float *base = lb_na_base(h->device);
    if (!base) return fail(PHC_ERR_STATE, "no scratch address for device %d", h->device);
    return lb_train_core(h, c, steps, batch_per_face, lr, seed, reinterpret_cast<int *>(base + LB_NA_IDX),
                         base + LB_NA_LOSS, base + LB_NA_PART, base + LB_NA_PLOSS, loss_history);
}

phc_status phc_lb_infer(phc_lb *h, int face, const float *samples, int64_t n, float *y)
{
    if (!h || !samples || !y) return fail(PHC_ERR_INVALID, "NULL argument");
    if (face < 0 || face >= LB_FACES) return fail(PHC_ERR_INVALID, "face %d outside 0..5", face);
    if (n < 1 || n > ((int64_t)1 << 30)) return fail(PHC_ERR_INVALID, "bad sample count");
    DevBuf s, out;
    PHC_TRY(s.alloc(4 * LB_NSTRIDE * (size_t)n));
    PHC_TRY(out.alloc(4 * (size_t)n));
    PHC_HIP(hipMemcpy(s.ptr, samples, 4 * LB_NSTRIDE * (size_t)n, hipMemcpyHostToDevice));
    const int64_t waves = (n + 31) / 32;
    hipLaunchKernelGGL(lb_infer_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, h->stream,
                       h->wbf.as<uint16_t>() + face * LB_NP, s.as<float>(), n, out.as<float>());
    PHC_HIP(hipGetLastError());
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemcpy(y, out.ptr, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return PHC_OK;
}

} // extern "C"

// bf16 parameters of a model, for an FD handle (fd.hip)
phc_status phc::lb_export_bf16(const phc_lb *h, uint16_t *dev_dst, hipStream_t stream)
{
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemcpyAsync(dev_dst, h->wbf.ptr, 2 * LB_FACES * LB_NP, hipMemcpyDeviceToDevice, stream));
    PHC_HIP(hipStreamSynchronize(stream));
    return PHC_OK;
}
