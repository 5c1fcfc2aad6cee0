// fd_tile.hpp -- the marching tile body of the FD step (inner3D + applySponge
// + fused source injection, src/fd/utilsFDTD/FDTDkernels.hpp:8-34, 108-128),
// shared by the single-grid step kernel (fd.hip) and the batched multi-shot
// kernel (batch.hip).
#pragma once

#include "fd_kernels.hpp"
#include "march.hpp"
#include "phc.h"

namespace phc {

// New value of one point, reference rounding: float lap accumulators, the
// update evaluated in double exactly as FDTDkernels.hpp:30-31 writes it, then
// the sponge product in float (FDTDkernels.hpp:114).
__device__ __forceinline__ float fd_update_ref(float pc, float pm, float v, double coef0,
                                               float lapx, float lapy, float lapz, float s)
{
#pragma clang fp contract(off)
    const double dpc = (double)pc;
    const double a = 2.0 * dpc - (double)pm;
    const double b = (((coef0 * dpc + (double)lapx) + (double)lapy) + (double)lapz);
    const float r = (float)(a + (double)v * b);
    return r * s;
}

// Face damping of one axis at index i of n (FDTDinit.hpp:343-395: the high
// face loop runs last, so it wins where the faces overlap).
__device__ __forceinline__ float fd_face_damping(const FdStepArgs &a, int i, int n)
{
    const int d = (i >= n - a.sw) ? a.sw - (n - i) : ((i < a.sw) ? a.sw - i : -1);
    return d >= 0 ? a.stab[d] : 1.f;
}

// One launch = inner3D + applySponge over the grid.
//
// A block owns a 64(k) x TJ(j) column and marches along i over chunk_i planes.
// Per thread and point: a ring of 2R+2 registers holds pn along i (planes
// i-R..i+R plus the one in flight); a ring of R+1 registers per halo element
// holds the j/k halo of the planes staged next; two register sets hold pnm1
// and vp^2dt^2 of planes i and i+1.  The march is unrolled by 2R+2, so every
// ring index is a compile-time constant and no value is ever moved between
// registers: a load has a full plane of work to complete before its register
// is read.  The halo of plane p is requested together with pn(p) of the
// register queue, i.e. when the neighbouring tile requests the same lines as
// its interior, so halo reads hit L2.  The current plane plus its halo is
// staged in LDS (double buffer, one LDS-only barrier per plane).
// IT is the offset type: 32-bit byte offsets when every array is < 4 GiB.
template <int R, int TY, int NJ>
struct FdTileShape {
    static constexpr int TJ = TY * NJ;
    static constexpr int LW = FD_TK + 2 * R;
    static constexpr int LH = TJ + 2 * R;
};

// March of one tile column (k-tile kt, j-tile jt, i-chunk ic) of one grid.
template <int R, int TY, int NJ, int ARITH, typename IT, class G>
__device__ __forceinline__ void fd_tile(const FdStepArgs &a, const G &g, int kt, int jt, int ic,
                                        float (*tile)[FdTileShape<R, TY, NJ>::LH][FdTileShape<R, TY, NJ>::LW])
{
    constexpr int TJ = TY * NJ;
    constexpr int LW = FD_TK + 2 * R;
    constexpr int NH = (2 * R + TY - 1) / TY;  // j-halo rows per thread
    constexpr int Q = 2 * R + 2;                     // queue ring = unroll factor
    constexpr int F = R + 1;                         // halo ring (Q == 2F)
    constexpr IT S = sizeof(IT) == 4 ? 4 : 1;        // offset units per element

    const FdLayout &L = g.L;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int k0 = kt * FD_TK, j0 = jt * TJ;
    const int i0 = g.ilo + ic * a.chunk_i;
    const int i1 = min(i0 + a.chunk_i, g.ihi);
    const int k = k0 + tx;
    const IT pitch = (IT)L.pitch_j * S;

    // Every load is unconditional: lanes or rows outside the stored grid read
    // the interior origin instead (their offsets point there), and what they
    // stage in LDS is never read by a point that is computed.
    const bool kload = k < L.nz + R;
    const bool kcomp = k >= g.klo && k < g.khi;
    int jr[NJ];
    bool jcomp[NJ], inJ[NJ];
    float fyv[NJ];
    IT colq[NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        jr[jj] = j0 + ty * NJ + jj;
        const bool jl = kload && (jr[jj] < L.ny + R);
        jcomp[jj] = kcomp && (jr[jj] >= g.jlo) && (jr[jj] < g.jhi);
        inJ[jj] = (jr[jj] >= a.sw) && (jr[jj] < L.ny - a.sw);
        fyv[jj] = (a.sw > 0 && jr[jj] < L.ny) ? g.fy[jr[jj]] : 1.f;
        colq[jj] = (IT)fd_off(L, 0, jl ? jr[jj] : 0, jl ? k : 0) * S;
    }
    const bool inK = (k >= a.sw) && (k < L.nz - a.sw);
    const float fzv = (a.sw > 0 && kcomp) ? g.fz[k] : 1.f;
    const bool khalo = tx < 2 * R;
    const int kk = (tx < R) ? (k0 - R + tx) : (k0 + FD_TK + tx - R);
    const int kcol = kk - k0 + R;
    IT colk[NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        const bool l = khalo && kk < L.nz + R && jr[jj] < L.ny + R;
        colk[jj] = (IT)fd_off(L, 0, l ? jr[jj] : 0, l ? kk : 0) * S;
    }
    int jhrow[NH];
    IT colh[NH];
#pragma unroll
    for (int m = 0; m < NH; ++m) {
        const int h = ty + TY * m;
        const int jh = (h < R) ? (j0 - R + h) : (j0 + TJ + h - R);
        jhrow[m] = jh - j0 + R;
        const bool l = (h < 2 * R) && (jh < L.ny + R) && kload;
        colh[m] = (IT)fd_off(L, 0, l ? jh : 0, l ? k : 0) * S;
    }

    float q[NJ][Q];                  // pn of plane p at slot (p - i0 + R) % Q
    float fk[NJ][F], fj[NH][F];      // halo of plane p at slot (p - i0) % F
    float pm[2][NJ], vv[2][NJ];      // pnm1 / vp of plane p in set (p - i0) % 2

    auto load_halo = [&](int slot, int p) {
        const IT po = (IT)p * pitch;
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) fk[jj][slot] = ldo(g.pn, colk[jj] + po);
#pragma unroll
        for (int m = 0; m < NH; ++m) fj[m][slot] = ldo(g.pn, colh[m] + po);
    };
    auto load_aux = [&](int set, int p) {
        const IT po = (IT)p * pitch;
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
            pm[set][jj] = ldo_nt(g.pnm1, colq[jj] + po);
            vv[set][jj] = ldo_nt(g.vp, colq[jj] + po);
        }
    };
    auto stage = [&](int buf, int qslot, int hslot) {
        float(*tl)[LW] = tile[buf];
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
            tl[ty * NJ + jj + R][tx + R] = q[jj][qslot];
            if (khalo) tl[ty * NJ + jj + R][kcol] = fk[jj][hslot];
        }
#pragma unroll
        for (int m = 0; m < NH; ++m)
            if (ty + TY * m < 2 * R) tl[jhrow[m]][tx + R] = fj[m][hslot];
    };

    // prologue: queue planes i0-R .. i0+R (slots 0..2R), halo planes
    // i0 .. i0+R (slots 0..R), pnm1/vp of plane i0 (set 0); stage plane i0
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj)
#pragma unroll
        for (int l = 0; l <= 2 * R; ++l) q[jj][l] = ldo(g.pn, colq[jj] + (IT)(i0 - R + l) * pitch);
#pragma unroll
    for (int d = 0; d <= R; ++d) load_halo(d, i0 + d);
    load_aux(0, i0);
    stage(0, R, 0);
    lds_barrier();

    for (int ib = i0; ib < i1; ib += Q) {
#pragma unroll
        for (int u = 0; u < Q; ++u) {
            const int i = ib + u;
            if (i >= i1) break;
            // Prefetch for plane i+1, issued unconditionally (past the end of
            // the grid the plane index is clamped; those values are never
            // used): with a conditional prefetch the wait-count insertion has
            // to assume no loads were issued and drains every load in flight.
            {
                const int pq = min(i + 1 + R, L.nx + R - 1);
                const IT po = (IT)pq * pitch;
#pragma unroll
                for (int jj = 0; jj < NJ; ++jj) q[jj][(u + 1 + 2 * R) % Q] = ldo(g.pn, colq[jj] + po);
                load_halo((u + 1 + R) % F, pq);
                load_aux((u + 1) & 1, min(i + 1, L.nx - 1));
            }
            // compute plane i
            const float(*tl)[LW] = tile[u & 1];
            const bool inI = (i >= a.sw) && (i < L.nx - a.sw);
            const float fxv = (a.sw > 0) ? fd_face_damping(a, i, L.nx) : 1.f;
            const IT pi = (IT)i * pitch;
#pragma unroll
            for (int jj = 0; jj < NJ; ++jj) {
                if (!jcomp[jj]) continue;
                const int r = ty * NJ + jj + R;
                const int c = tx + R;
                float s = 1.f;
                if (a.sw > 0) {
                    if (inJ[jj] && inK && !inI) s = fxv;
                    if (inK && inI && !inJ[jj]) s = fyv[jj];
                    if (inJ[jj] && inI && !inK) s = fzv;
                }
                const float pc = q[jj][(u + R) % Q];
                const float pmj = pm[u & 1][jj], vvj = vv[u & 1][jj];
                float val;
                if constexpr (ARITH == PHC_FD_ARITH_REF) {
                    float lapx, lapy, lapz;
                    {
#pragma clang fp contract(off)
                        lapx = 0.f;
#pragma unroll
                        for (int l = 1; l <= R; ++l)
                            lapx += a.cx[l] * (q[jj][(u + R + l) % Q] + q[jj][(u + R - l) % Q]);
                        lapy = 0.f;
#pragma unroll
                        for (int l = 1; l <= R; ++l) lapy += a.cy[l] * (tl[r + l][c] + tl[r - l][c]);
                        lapz = 0.f;
#pragma unroll
                        for (int l = 1; l <= R; ++l) lapz += a.cz[l] * (tl[r][c + l] + tl[r][c - l]);
                    }
                    // deferred damping of the previous level (x1 is exact)
                    const float pmv = mul_nc(pmj, a.damp_prev ? s : 1.f);
                    val = fd_update_ref(pc, pmv, vvj, a.coef0, lapx, lapy, lapz, s);
                } else {
                    float lap = a.c0f * pc;
#pragma unroll
                    for (int l = 1; l <= R; ++l) {
                        lap = fmaf(a.cx[l], q[jj][(u + R + l) % Q] + q[jj][(u + R - l) % Q], lap);
                        lap = fmaf(a.cy[l], tl[r + l][c] + tl[r - l][c], lap);
                        lap = fmaf(a.cz[l], tl[r][c + l] + tl[r][c - l], lap);
                    }
                    const float pmv = pmj * (a.damp_prev ? s : 1.f);
                    val = s * fmaf(vvj, lap, 2.f * pc - pmv);
                }
                const IT o = colq[jj] + pi;
                if ((int64_t)(o / S) == g.src_off) {
                    if (g.trace) *g.trace = val;
                    if (g.inj) val = add_nc(val, mul_nc(vvj, g.rhs_next));
                }
                sto_nt(g.pnm1, o, val);
            }
            stage((u + 1) & 1, (u + 1 + R) % Q, (u + 1) % F);
            lds_barrier();
        }
    }
}

} // namespace phc
