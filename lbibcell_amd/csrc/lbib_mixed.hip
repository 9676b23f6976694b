// lbib_mixed.hip -- mixed-precision lattice kernels: FP32 arithmetic on
// populations stored as FP32, FP16 or BF16 deviations from the rest state
// (f_i - w_i for the fluid; species populations are stored as they are,
// their rest state being C = 0).  Membrane markers, forces and areas stay
// FP64; the velocity field is kept as FP32 pairs.
//
// The physics is that of the FP64 kernels (FluidSolver::collide,
// FluidSolver.cpp:103-183 with Luo1993ForceModel.cpp:29-70 or
// GuoZhengShi2002ForceModel.cpp:29-78; CDESolverD2Q5::collide,
// CDESolverD2Q5.cpp:75-107 and the SIGNAL reaction; the no-flux boundary of
// BoundarySolverNoFluxD2Q5.cpp:37-113; MassSolverBoxOutlet.cpp:42-72), written
// in deviation form so that the rest state cancels exactly:
//   rho0 = 1 + sum(d),  u = (momentum of d [+ F/2]) / rho,
//   d_eq = w (rho - 1) + w rho (3 e.u + 9/2 (e.u)^2 - 3/2 u^2),
//   d'   = d - (d - d_eq) / tau + force term.
// Rounding differs from the reference by construction; the error against
// the FP64 path is measured, not bounded here.
//
// Roofline: HBM bound.  Algorithmic bytes per site = (2*9 + 2*5*ns) * b with
// b = 4 (FP32) or 2 (FP16/BF16).
#include "lbib_mixed.h"

namespace lbib {

// SPT sites per thread, one workgroup width (blockDim.x) apart so that every
// load stays coalesced; all loads of a thread are issued before its
// arithmetic (16-bit values carry few bytes per instruction, so the bytes in
// flight per wavefront matter).  SPT and the width are the tuned launch shape.
template <int NS, int FM, int P, int SPT>
__global__ __launch_bounds__(256) void k_stream_collide_mx(const MixedArgs a) {
    using C = Codec<P>;
    using T = typename C::T;
    if (step_stopped(a.st)) return;
    const int W = (int)blockDim.x;
    const uint32_t wg = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int x0w = (int)(wg % gridDim.x) * (W * SPT);   // the workgroup's first site
    const int xb = x0w + threadIdx.x;
    const int y = (int)(wg / gridDim.x);
    if (xb >= a.nx) return;
    const long n = a.n;
    const long r0 = (long)y * a.nx;
    const T* fs = static_cast<const T*>(a.fs);
    const T* gsb = static_cast<const T*>(a.gs);
    constexpr int GQ = NS > 0 ? 5 * NS : 5;   // species planes in the storage blocks

    bool live[SPT];
    int xs[SPT];
    Frame fr[SPT];
    float d[SPT][9];
    float g[SPT][NS > 0 ? NS : 1][5];
    uint32_t tg[SPT];
    float Fx[SPT], Fy[SPT], m[SPT], pu[SPT], pv[SPT];
    bool any = false;
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        const int x = xb + k * W;
        live[k] = x < a.nx && site_frame(a.rtile, a.rects, a.ntx, a.nx, a.ny, x, y, fr[k]);
        if (!live[k]) { fr[k].x0 = 0; fr[k].y0 = 0; fr[k].w = a.nx; fr[k].h = a.ny; fr[k].r = -1; }
        // a site outside every rectangle of a lane that has live sites reads
        // the workgroup's first site, one address for all such lanes, and
        // writes nothing
        xs[k] = live[k] ? x : x0w;
        any |= live[k];
    }
    // a lane whose sites all lie outside every rectangle has nothing to do
    // (no workgroup barrier or cross-lane operation follows)
    if (!any) return;
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        const int xx = xs[k];
        const int xm = (xx == fr[k].x0) ? fr[k].x0 + fr[k].w - 1 : xx - 1;
        const int xp = (xx == fr[k].x0 + fr[k].w - 1) ? fr[k].x0 : xx + 1;
        const long rS = (long)((y == fr[k].y0) ? fr[k].y0 + fr[k].h - 1 : y - 1) * a.nx;
        const long rN = (long)((y == fr[k].y0 + fr[k].h - 1) ? fr[k].y0 : y + 1) * a.nx;
        const long i = r0 + xx;
        // pulled deviations d_q = f_q - w_q (each read once per step: non-temporal)
        d[k][DT] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DT, i, n, 9)]));
        d[k][DE] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DE, r0 + xm, n, 9)]));
        d[k][DN] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DN, rS + xx, n, 9)]));
        d[k][DW] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DW, r0 + xp, n, 9)]));
        d[k][DS] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DS, rN + xx, n, 9)]));
        d[k][DNE] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DNE, rS + xm, n, 9)]));
        d[k][DNW] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DNW, rS + xp, n, 9)]));
        d[k][DSW] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DSW, rN + xp, n, 9)]));
        d[k][DSE] = C::dec(__builtin_nontemporal_load(&fs[pidx<P>(DSE, rN + xm, n, 9)]));
#pragma unroll
        for (int s = 0; s < NS; s++) {
            g[k][s][DT] = C::dec(__builtin_nontemporal_load(&gsb[pidx<P>(5 * s + DT, i, n, GQ)]));
            g[k][s][DE] = C::dec(__builtin_nontemporal_load(&gsb[pidx<P>(5 * s + DE, r0 + xm, n, GQ)]));
            g[k][s][DN] = C::dec(__builtin_nontemporal_load(&gsb[pidx<P>(5 * s + DN, rS + xx, n, GQ)]));
            g[k][s][DW] = C::dec(__builtin_nontemporal_load(&gsb[pidx<P>(5 * s + DW, r0 + xp, n, GQ)]));
            g[k][s][DS] = C::dec(__builtin_nontemporal_load(&gsb[pidx<P>(5 * s + DS, rN + xx, n, GQ)]));
        }
        tg[k] = a.tag[i];
        Fx[k] = 0.0f;
        Fy[k] = 0.0f;
        if (a.ftile == nullptr || a.ftile[(long)(y >> 4) * a.ntx + (xx >> 4)]) {
            const float2 F = a.F[i];
            Fx[k] = F.x;
            Fy[k] = F.y;
        }
    }
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        const long i = r0 + xs[k];
        const uint32_t dom = tg[k] & DOM_MASK;
        m[k] = 0.0f;
        if (a.msite) { m[k] = (float)a.msite[i]; if (live[k]) a.msite[i] = 0.0; }
        else if (a.mtab) m[k] = (dom < (uint32_t)a.ntab) ? (float)a.mtab[dom] : 0.0f;
        pu[k] = 0.0f;
        pv[k] = 0.0f;
        if (FM != 0 || (NS > 0 && (tg[k] >> BC_SHIFT))) {
            const float2 up = *reinterpret_cast<const float2*>(a.u + 2 * i);
            pu[k] = up.x;
            pv[k] = up.y;
        }
        if (!C::dev) {
#pragma unroll
            for (int q = 0; q < 9; q++) d[k][q] -= wq(q);
        }
    }

    T* fd = static_cast<T*>(a.fd);
    T* gd = static_cast<T*>(a.gd);
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        if (!live[k]) continue;
        const int x = xs[k];
        const long i = r0 + x;
        const uint32_t dom = tg[k] & DOM_MASK;
        if (NS > 0 && (tg[k] >> BC_SHIFT)) {
            // no-flux cut links: incoming population from the first listed
            // species' pre-stream state, swapped into every listed species
            const int inv[5] = {DT, DW, DS, DE, DN};
            for (int o = 1; o <= 4; o++) {
                if (!((tg[k] >> (BC_SHIFT + o - 1)) & 1u)) continue;
                const int dd = inv[o];
                const long e = (long)(o - 1) * n + i;
                const int desc = a.bc_desc[e];
                const int b0 = a.bcd_off[desc], b1 = a.bcd_off[desc + 1];
                if (b1 <= b0) continue;
                const int s1 = 5 * a.bcd_sp[b0];
                const float Cp = C::dec(gsb[pidx<P>(s1, i, n, GQ)]) + C::dec(gsb[pidx<P>(s1 + 1, i, n, GQ)]) +
                                 C::dec(gsb[pidx<P>(s1 + 2, i, n, GQ)]) + C::dec(gsb[pidx<P>(s1 + 3, i, n, GQ)]) +
                                 C::dec(gsb[pidx<P>(s1 + 4, i, n, GQ)]);
                const float c6 = Cp / 6.0f;
                const float bvx = (float)a.bc_v[2 * e], bvy = (float)a.bc_v[2 * e + 1];
                float feq, eq;
                if (dd == DE) { feq = c6 * (1.0f + 3.0f * bvx); eq = c6 * (1.0f + 3.0f * pu[k]); }
                else if (dd == DN) { feq = c6 * (1.0f + 3.0f * bvy); eq = c6 * (1.0f + 3.0f * pv[k]); }
                else if (dd == DW) { feq = c6 * (1.0f - 3.0f * bvx); eq = c6 * (1.0f - 3.0f * pu[k]); }
                else { feq = c6 * (1.0f - 3.0f * bvy); eq = c6 * (1.0f - 3.0f * pv[k]); }
                float carry = feq + (C::dec(gsb[pidx<P>(s1 + dd, i, n, GQ)]) - eq);
                for (int kk = b0; kk < b1; kk++) {
                    const int sk = a.bcd_sp[kk];
#pragma unroll
                    for (int s = 0; s < NS; s++)
                        if (s == sk) { const float t = g[k][s][dd]; g[k][s][dd] = carry; carry = t; }
                }
            }
        }

        float uxn, uyn;
        const float tauI = fr[k].r < 0 ? a.tauI : (float)a.rects[fr[k].r].tauI;
        const float tau = fr[k].r < 0 ? a.tau : (float)a.rects[fr[k].r].tau;
        collide_dev_f32<FM>(d[k], m[k], Fx[k], Fy[k], tauI, tau, pu[k], pv[k], uxn, uyn);

        if (NS > 0) {
            float cj[NS > 0 ? NS : 1], cseq[NS > 0 ? NS : 1];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                cj[s] = g[k][s][0] + g[k][s][1] + g[k][s][2] + g[k][s][3] + g[k][s][4];
                cseq[s] = cj[s];
            }
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const float tI = fr[k].r < 0 ? a.sp[s].tauI : (float)a.rects[fr[k].r].tauIs[s];
                species_collide_f32<NS>(g[k][s], cj[s], uxn, uyn, a.sp[s], tI, dom, iter_of(a.it, a.st), cj, cseq);
                // a later species sees this one's concentration after its collide
                cseq[s] = g[k][s][0] + g[k][s][1] + g[k][s][2] + g[k][s][3] + g[k][s][4];
            }
        }

        if (a.box_outlet && (x == fr[k].x0 || y == fr[k].y0)) {
            const int passes = (x == fr[k].x0) + (y == fr[k].y0);
            for (int p = 0; p < passes; p++) {
                float sd = 0.0f;
#pragma unroll
                for (int q = 0; q < 9; q++) sd += d[k][q];
                const float sf = 1.0f / (1.0f + sd);
#pragma unroll
                for (int q = 0; q < 9; q++) d[k][q] = (d[k][q] - wq(q) * sd) * sf;   // (w + d) / (1 + sd) - w
#pragma unroll
                for (int s = 0; s < NS; s++)
#pragma unroll
                    for (int q = 0; q < 5; q++) g[k][s][q] *= sf;
            }
        }

#pragma unroll
        for (int q = 0; q < 9; q++)
            __builtin_nontemporal_store(C::enc(C::dev ? d[k][q] : d[k][q] + wq(q)), &fd[pidx<P>(q, i, n, 9)]);
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++)
                __builtin_nontemporal_store(C::enc(g[k][s][q]), &gd[pidx<P>(5 * s + q, i, n, GQ)]);
        *reinterpret_cast<float2*>(a.u + 2 * i) = make_float2(uxn, uyn);
    }
}

// 16-bit storage, even nx: each lane owns the adjacent sites (x0, x0+1), so
// every load and store moves 32 bits per lane.  Pulls shifted by one site
// along x combine the lane's aligned pair with its neighbour lane's (a
// wavefront shuffle); the lanes at the wavefront's ends (and at the periodic
// wrap) load the one value from outside the wavefront themselves.
template <int P>
struct PairLoad {
    using T = uint16_t;
    const T* base;
    long r, n;               // row start site, sites
    int q, nq;               // plane, planes per storage block
    int x0, fx0, fw, lane;   // pair start; periodic frame [fx0, fx0 + fw) along x
    bool lo_edge, hi_edge;
    __device__ long at(int x) const { return pidx<P>(q, r + x, n, nq); }
    // the pair (x0, x0 + 1) is contiguous: x0 and the row start are even
    // (each population is read once per step: non-temporal load)
    __device__ uint32_t aligned() const {
        return __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(base + at(x0)));
    }
    // values at x0-1, x0 (pull from the west)
    __device__ void west(float& a0, float& a1) const {
        const uint32_t own = aligned();
        uint32_t prev = (uint32_t)__shfl_up((int)own, 1, 64);
        const T lo = lo_edge ? base[at(x0 == fx0 ? fx0 + fw - 1 : x0 - 1)] : (T)(prev >> 16);
        a0 = Codec<P>::dec(lo);
        a1 = Codec<P>::dec((T)(own & 0xffffu));
    }
    // values at x0+1, x0+2 (pull from the east)
    __device__ void east(float& a0, float& a1) const {
        const uint32_t own = aligned();
        uint32_t next = (uint32_t)__shfl_down((int)own, 1, 64);
        const T hi = hi_edge ? base[at(x0 + 2 >= fx0 + fw ? x0 + 2 - fw : x0 + 2)] : (T)(next & 0xffffu);
        a0 = Codec<P>::dec((T)(own >> 16));
        a1 = Codec<P>::dec(hi);
    }
    __device__ void here(float& a0, float& a1) const {
        const uint32_t own = aligned();
        a0 = Codec<P>::dec((T)(own & 0xffffu));
        a1 = Codec<P>::dec((T)(own >> 16));
    }
};

template <int NS, int FM, int P>
__global__ __launch_bounds__(256) void k_stream_collide_h16(const MixedArgs a) {
    using C = Codec<P>;
    using T = uint16_t;
    if (step_stopped(a.st)) return;
    const int W = (int)blockDim.x;
    const uint32_t wg = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int lane = threadIdx.x & 63;
    const int x0 = (int)(wg % gridDim.x) * (2 * W) + 2 * (int)threadIdx.x;
    const int y = (int)(wg / gridDim.x);
    // lanes past the row end or outside every domain take part in the
    // shuffles with a valid address; both sites of a pair share one 16x16 tile
    Frame fr;
    const bool live = x0 < a.nx && site_frame(a.rtile, a.rects, a.ntx, a.nx, a.ny, x0, y, fr);
    if (!live) { fr.x0 = 0; fr.y0 = 0; fr.w = a.nx; fr.h = a.ny; fr.r = -1; }
    const int xl = x0 < a.nx ? x0 : a.nx - 2;
    const long n = a.n;
    const long r0 = (long)y * a.nx;
    const long rS = (long)((y == fr.y0) ? fr.y0 + fr.h - 1 : y - 1) * a.nx;
    const long rN = (long)((y == fr.y0 + fr.h - 1) ? fr.y0 : y + 1) * a.nx;
    const T* fs = static_cast<const T*>(a.fs);
    const T* gsb = static_cast<const T*>(a.gs);
    constexpr int GQ = NS > 0 ? 5 * NS : 5;   // species planes in the storage blocks
    const bool lo_edge = lane == 0 || xl == fr.x0;
    const bool hi_edge = lane == 63 || xl + 2 >= fr.x0 + fr.w;
    auto pl = [&](const T* base, int q, int nq, long row) {
        return PairLoad<P>{base, row, n, q, nq, xl, fr.x0, fr.w, lane, lo_edge, hi_edge};
    };

    float d[2][9];
    pl(fs, DT, 9, r0).here(d[0][DT], d[1][DT]);
    pl(fs, DE, 9, r0).west(d[0][DE], d[1][DE]);
    pl(fs, DN, 9, rS).here(d[0][DN], d[1][DN]);
    pl(fs, DW, 9, r0).east(d[0][DW], d[1][DW]);
    pl(fs, DS, 9, rN).here(d[0][DS], d[1][DS]);
    pl(fs, DNE, 9, rS).west(d[0][DNE], d[1][DNE]);
    pl(fs, DNW, 9, rS).east(d[0][DNW], d[1][DNW]);
    pl(fs, DSW, 9, rN).east(d[0][DSW], d[1][DSW]);
    pl(fs, DSE, 9, rN).west(d[0][DSE], d[1][DSE]);
    float g[2][NS > 0 ? NS : 1][5];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        pl(gsb, 5 * s + DT, GQ, r0).here(g[0][s][DT], g[1][s][DT]);
        pl(gsb, 5 * s + DE, GQ, r0).west(g[0][s][DE], g[1][s][DE]);
        pl(gsb, 5 * s + DN, GQ, rS).here(g[0][s][DN], g[1][s][DN]);
        pl(gsb, 5 * s + DW, GQ, r0).east(g[0][s][DW], g[1][s][DW]);
        pl(gsb, 5 * s + DS, GQ, rN).here(g[0][s][DS], g[1][s][DS]);
    }
    if (!live) return;
    const long i0 = r0 + x0;
    const uint2 tg2 = *reinterpret_cast<const uint2*>(a.tag + i0);
    const uint32_t tg[2] = {tg2.x, tg2.y};
    float Fx[2] = {0.0f, 0.0f}, Fy[2] = {0.0f, 0.0f};
    if (a.ftile == nullptr || a.ftile[(long)(y >> 4) * a.ntx + (x0 >> 4)]) {
        const float4 F = *reinterpret_cast<const float4*>(a.F + i0);
        Fx[0] = F.x; Fy[0] = F.y; Fx[1] = F.z; Fy[1] = F.w;
    }
    float pu[2] = {0.0f, 0.0f}, pv[2] = {0.0f, 0.0f};
    if (FM != 0 || (NS > 0 && ((tg[0] | tg[1]) >> BC_SHIFT))) {
        const float4 up = *reinterpret_cast<const float4*>(a.u + 2 * i0);
        pu[0] = up.x; pv[0] = up.y; pu[1] = up.z; pv[1] = up.w;
    }
    float uo[4];
    uint32_t outf[9] = {}, outg[NS > 0 ? NS : 1][5] = {};
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int x = x0 + k;
        const long i = i0 + k;
        const uint32_t dom = tg[k] & DOM_MASK;
        float m = 0.0f;
        if (a.msite) { m = (float)a.msite[i]; a.msite[i] = 0.0; }
        else if (a.mtab) m = (dom < (uint32_t)a.ntab) ? (float)a.mtab[dom] : 0.0f;
        if (NS > 0 && (tg[k] >> BC_SHIFT)) {
            const int inv[5] = {DT, DW, DS, DE, DN};
            for (int o = 1; o <= 4; o++) {
                if (!((tg[k] >> (BC_SHIFT + o - 1)) & 1u)) continue;
                const int dd = inv[o];
                const long e = (long)(o - 1) * n + i;
                const int desc = a.bc_desc[e];
                const int b0 = a.bcd_off[desc], b1 = a.bcd_off[desc + 1];
                if (b1 <= b0) continue;
                const int s1 = 5 * a.bcd_sp[b0];
                const float Cp = C::dec(gsb[pidx<P>(s1, i, n, GQ)]) + C::dec(gsb[pidx<P>(s1 + 1, i, n, GQ)]) +
                                 C::dec(gsb[pidx<P>(s1 + 2, i, n, GQ)]) + C::dec(gsb[pidx<P>(s1 + 3, i, n, GQ)]) +
                                 C::dec(gsb[pidx<P>(s1 + 4, i, n, GQ)]);
                const float c6 = Cp / 6.0f;
                const float bvx = (float)a.bc_v[2 * e], bvy = (float)a.bc_v[2 * e + 1];
                float feq, eq;
                if (dd == DE) { feq = c6 * (1.0f + 3.0f * bvx); eq = c6 * (1.0f + 3.0f * pu[k]); }
                else if (dd == DN) { feq = c6 * (1.0f + 3.0f * bvy); eq = c6 * (1.0f + 3.0f * pv[k]); }
                else if (dd == DW) { feq = c6 * (1.0f - 3.0f * bvx); eq = c6 * (1.0f - 3.0f * pu[k]); }
                else { feq = c6 * (1.0f - 3.0f * bvy); eq = c6 * (1.0f - 3.0f * pv[k]); }
                float carry = feq + (C::dec(gsb[pidx<P>(s1 + dd, i, n, GQ)]) - eq);
                for (int kk = b0; kk < b1; kk++) {
                    const int sk = a.bcd_sp[kk];
#pragma unroll
                    for (int s = 0; s < NS; s++)
                        if (s == sk) { const float t = g[k][s][dd]; g[k][s][dd] = carry; carry = t; }
                }
            }
        }
        float uxn, uyn;
        const float tauI = fr.r < 0 ? a.tauI : (float)a.rects[fr.r].tauI;
        const float tau = fr.r < 0 ? a.tau : (float)a.rects[fr.r].tau;
        collide_dev_f32<FM>(d[k], m, Fx[k], Fy[k], tauI, tau, pu[k], pv[k], uxn, uyn);
        if (NS > 0) {
            float cj[NS > 0 ? NS : 1], cseq[NS > 0 ? NS : 1];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                cj[s] = g[k][s][0] + g[k][s][1] + g[k][s][2] + g[k][s][3] + g[k][s][4];
                cseq[s] = cj[s];
            }
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const float tI = fr.r < 0 ? a.sp[s].tauI : (float)a.rects[fr.r].tauIs[s];
                species_collide_f32<NS>(g[k][s], cj[s], uxn, uyn, a.sp[s], tI, dom, iter_of(a.it, a.st), cj, cseq);
                // a later species sees this one's concentration after its collide
                cseq[s] = g[k][s][0] + g[k][s][1] + g[k][s][2] + g[k][s][3] + g[k][s][4];
            }
        }
        if (a.box_outlet && (x == fr.x0 || y == fr.y0)) {
            const int passes = (x == fr.x0) + (y == fr.y0);
            for (int p = 0; p < passes; p++) {
                float sd = 0.0f;
#pragma unroll
                for (int q = 0; q < 9; q++) sd += d[k][q];
                const float sf = 1.0f / (1.0f + sd);
#pragma unroll
                for (int q = 0; q < 9; q++) d[k][q] = (d[k][q] - wq(q) * sd) * sf;   // (w + d) / (1 + sd) - w
#pragma unroll
                for (int s = 0; s < NS; s++)
#pragma unroll
                    for (int q = 0; q < 5; q++) g[k][s][q] *= sf;
            }
        }
#pragma unroll
        for (int q = 0; q < 9; q++) outf[q] |= k ? ((uint32_t)C::enc(d[k][q]) << 16) : (uint32_t)C::enc(d[k][q]);
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++)
                outg[s][q] |= k ? ((uint32_t)C::enc(g[k][s][q]) << 16) : (uint32_t)C::enc(g[k][s][q]);
        uo[2 * k] = uxn;
        uo[2 * k + 1] = uyn;
    }
    T* fd = static_cast<T*>(a.fd);
    T* gd = static_cast<T*>(a.gd);
#pragma unroll
    for (int q = 0; q < 9; q++)
        __builtin_nontemporal_store(outf[q], reinterpret_cast<uint32_t*>(fd + pidx<P>(q, i0, n, 9)));
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 5; q++)
            __builtin_nontemporal_store(outg[s][q], reinterpret_cast<uint32_t*>(gd + pidx<P>(5 * s + q, i0, n, GQ)));
    *reinterpret_cast<float4*>(a.u + 2 * i0) = make_float4(uo[0], uo[1], uo[2], uo[3]);
}

// Vector lattice kernel.  Each lane owns V consecutive sites of one row, so
// one population plane moves in one 8- or 16-byte load and one store per
// lane.  A pull shifted along x takes the missing site from the neighbouring
// lane with one wavefront shuffle per plane; lanes 0 and 63 of a wavefront
// and the lanes holding a periodic frame's first or last site load that one
// site themselves (all such loads in two branch regions, after every vector
// load has been issued).
//
// A workgroup serves one row segment [xs, xe) of one periodic frame: a chunk
// of a lattice row, or in a batched atlas a chunk of a row of one rectangle
// taken from a segment table, so the frame and its relaxation times are
// uniform over the workgroup and no lane is launched outside a rectangle.
// Rectangle origins are multiples of 16, so a frame always starts at a lane's
// first site; a frame may end inside a lane (odd multiple of 2 widths), whose
// sites past the end are neither stored nor used.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <int P, int V>
struct LaneVec {
    using C = Codec<P>;
    using T = typename C::T;
    static constexpr int NW = V * (int)sizeof(T) / 4;   // dwords per lane
    uint32_t w[NW];
    uint32_t edge;   // storage bits of the site shifted in from outside the lane
    // every population is read once per step: non-temporal loads
    __device__ void load(const T* p) {
        const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
        if constexpr (NW % 4 == 0) {
#pragma unroll
            for (int k = 0; k < NW; k += 4) {
                const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(q + k));
                w[k] = v.x; w[k + 1] = v.y; w[k + 2] = v.z; w[k + 3] = v.w;
            }
        } else {
            static_assert(NW == 2, "lane vectors of 8 or a multiple of 16 bytes");
            const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(q));
            w[0] = v.x; w[1] = v.y;
        }
    }
    __device__ uint32_t bits(int j) const {
        if constexpr (sizeof(T) == 2) return (w[j >> 1] >> (16 * (j & 1))) & 0xffffu;
        else return w[j];
    }
    static __device__ float dec(uint32_t b) {
        if constexpr (sizeof(T) == 2) return C::dec((T)b);
        else return __uint_as_float(b);
    }
    static __device__ uint32_t enc(float x) {
        if constexpr (sizeof(T) == 2) return (uint32_t)C::enc(x);
        else return __float_as_uint(x);
    }
    static __device__ uint32_t ld1(const T* p) {
        if constexpr (sizeof(T) == 2) return (uint32_t)*reinterpret_cast<const uint16_t*>(p);
        else return *reinterpret_cast<const uint32_t*>(p);
    }
    // site j of the plane pulled with shift s (0 here, 1 from the west, 2
    // from the east; je: the lane's last site whose east neighbour is edge)
    __device__ float pulled(int s, int j, int je) const {
        if (s == 0) return dec(bits(j));
        if (s == 1) return dec(j == 0 ? edge : bits(j - 1));
        if (j == V - 1) return dec(edge);
        return dec(j >= je ? edge : bits(j + 1));
    }
    __device__ void store(T* p) const {
        uint32_t* q = reinterpret_cast<uint32_t*>(p);
        if constexpr (NW % 4 == 0) {
#pragma unroll
            for (int k = 0; k < NW; k += 4)
                __builtin_nontemporal_store(u32x4{w[k], w[k + 1], w[k + 2], w[k + 3]},
                                            reinterpret_cast<u32x4*>(q + k));
        } else {
            __builtin_nontemporal_store(u32x2{w[0], w[1]}, reinterpret_cast<u32x2*>(q));
        }
    }
    __device__ void store_first(T* p, int count) const {   // the first count sites
#pragma unroll
        for (int j = 0; j < V; j++) {
            if (j >= count) break;
            if constexpr (sizeof(T) == 2) reinterpret_cast<uint16_t*>(p)[j] = (uint16_t)bits(j);
            else reinterpret_cast<uint32_t*>(p)[j] = bits(j);
        }
    }
    __device__ void put(int j, uint32_t b) {
        if constexpr (sizeof(T) == 2) {
            if (j & 1) w[j >> 1] |= b << 16;
            else w[j >> 1] = b;
        } else {
            w[j] = b;
        }
    }
};

// A batched atlas for the vector kernel: rectangle r's sites as groups of V
// consecutive sites along its rows, numbered row-major (a group never crosses
// a row); one workgroup takes groups l0 .. l0 + blockDim - 1, so a workgroup
// may cover the ends of several short rows and only a rectangle's last
// workgroup has lanes without sites.
struct SegDev {
    int32_t r, l0;
};

// plane -> source row (0 same, 1 south, 2 north) and x shift (0, 1 west, 2 east)
__device__ constexpr int row_of(int q) {
    return (q == DN || q == DNE || q == DNW) ? 1 : ((q == DS || q == DSW || q == DSE) ? 2 : 0);
}
__device__ constexpr int shift_of(int q) {
    return (q == DE || q == DNE || q == DSE) ? 1 : ((q == DW || q == DNW || q == DSW) ? 2 : 0);
}

template <int NS, int FM, int P, int V>
__global__ __launch_bounds__(256) void k_stream_collide_vec(const MixedArgs a, const SegDev* __restrict__ segs,
                                                            int nchunk) {
    using LV = LaneVec<P, V>;
    using T = typename LV::T;
    if (step_stopped(a.st)) return;
    const uint32_t wg = xcd_contiguous(blockIdx.x, gridDim.x);
    int y, x, nvalid, fx0, fy0, fw, fh, r;
    if (segs) {
        const SegDev sg = segs[wg];
        r = sg.r;
        const RectDev& R = a.rects[r];
        fx0 = R.x0; fy0 = R.y0; fw = R.w; fh = R.h;
        const int G = (fw + V - 1) / V;   // groups per row
        const int l = sg.l0 + (int)threadIdx.x;
        const bool dead = l >= G * fh;
        const int row = dead ? 0 : l / G;
        y = fy0 + row;
        x = dead ? fx0 : fx0 + V * (l - row * G);
        nvalid = dead ? 0 : min(fx0 + fw - x, V);
    } else {
        const int L = (int)blockDim.x * V;
        y = (int)(wg / (uint32_t)nchunk);
        const int xs = (int)(wg % (uint32_t)nchunk) * L;
        x = xs + V * (int)threadIdx.x;
        nvalid = min(max(min(a.nx, xs + L) - x, 0), V);
        r = -1; fx0 = 0; fy0 = 0; fw = a.nx; fh = a.ny;
    }
    const int fxe = fx0 + fw;
    const int lane = (int)threadIdx.x & 63;
    // dead lanes keep a valid address; a live group lies inside the lattice
    // (origins on the 16-site grid, even widths, nx a multiple of 16)
    const int xl = min(x, a.nx - V);
    const long n = a.n;
    const long r0 = (long)y * a.nx;
    const long rS = (long)((y == fy0) ? fy0 + fh - 1 : y - 1) * a.nx;
    const long rN = (long)((y == fy0 + fh - 1) ? fy0 : y + 1) * a.nx;
    auto rsel = [&](int k) { return k == 0 ? r0 : (k == 1 ? rS : rN); };
    const bool w_need = lane == 0 || xl == fx0;
    const int w_x = (xl == fx0) ? fxe - 1 : xl - 1;
    const bool e_fend = xl + V - 1 >= fxe - 1;
    const bool e_need = lane == 63 || e_fend;
    const int e_x = e_fend ? fx0 : xl + V;
    const int je = e_fend ? min(max(fxe - 1 - xl, 0), V - 1) : V - 1;
    const T* fs = static_cast<const T*>(a.fs);
    const T* gsb = static_cast<const T*>(a.gs);
    constexpr int NSS = NS > 0 ? NS : 1;

    LV f[9], g[NSS][5];
#pragma unroll
    for (int q = 0; q < 9; q++) f[q].load(fs + (long)q * n + rsel(row_of(q)) + xl);
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) g[s][q].load(gsb + (long)(5 * s + q) * n + rsel(row_of(q)) + xl);
    const long i0 = r0 + xl;
    uint32_t tg[V];
#pragma unroll
    for (int k = 0; k < V; k += 2) {
        const uint2 t = *reinterpret_cast<const uint2*>(a.tag + i0 + k);
        tg[k] = t.x; tg[k + 1] = t.y;
    }
    float Fx[V], Fy[V];
#pragma unroll
    for (int k = 0; k < V; k++) { Fx[k] = 0.0f; Fy[k] = 0.0f; }
    if (a.ftile == nullptr || a.ftile[(long)(y >> 4) * a.ntx + (xl >> 4)]) {
#pragma unroll
        for (int k = 0; k < V; k += 2) {
            const float4 F = *reinterpret_cast<const float4*>(a.F + i0 + k);
            Fx[k] = F.x; Fy[k] = F.y; Fx[k + 1] = F.z; Fy[k + 1] = F.w;
        }
    }
    // sites shifted in from outside the lane
    uint32_t wh[9 + 5 * NSS], eh[9 + 5 * NSS];
    if (w_need) {
#pragma unroll
        for (int q = 0; q < 9; q++)
            if (shift_of(q) == 1) wh[q] = LV::ld1(fs + (long)q * n + rsel(row_of(q)) + w_x);
#pragma unroll
        for (int s = 0; s < NS; s++)
            wh[9 + 5 * s + DE] = LV::ld1(gsb + (long)(5 * s + DE) * n + r0 + w_x);
    }
    if (e_need) {
#pragma unroll
        for (int q = 0; q < 9; q++)
            if (shift_of(q) == 2) eh[q] = LV::ld1(fs + (long)q * n + rsel(row_of(q)) + e_x);
#pragma unroll
        for (int s = 0; s < NS; s++)
            eh[9 + 5 * s + DW] = LV::ld1(gsb + (long)(5 * s + DW) * n + r0 + e_x);
    }
#pragma unroll
    for (int q = 0; q < 9; q++) {
        if (shift_of(q) == 1) {
            const uint32_t sh = (uint32_t)__shfl_up((int)f[q].bits(V - 1), 1, 64);
            f[q].edge = w_need ? wh[q] : sh;
        } else if (shift_of(q) == 2) {
            const uint32_t sh = (uint32_t)__shfl_down((int)f[q].bits(0), 1, 64);
            f[q].edge = e_need ? eh[q] : sh;
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const uint32_t shw = (uint32_t)__shfl_up((int)g[s][DE].bits(V - 1), 1, 64);
        g[s][DE].edge = w_need ? wh[9 + 5 * s + DE] : shw;
        const uint32_t she = (uint32_t)__shfl_down((int)g[s][DW].bits(0), 1, 64);
        g[s][DW].edge = e_need ? eh[9 + 5 * s + DW] : she;
    }

    bool anybc = false;
#pragma unroll
    for (int k = 0; k < V; k++) anybc |= (tg[k] >> BC_SHIFT) != 0u;
    float pu[V], pv[V];
#pragma unroll
    for (int k = 0; k < V; k++) { pu[k] = 0.0f; pv[k] = 0.0f; }
    if (FM != 0 || (NS > 0 && anybc)) {
#pragma unroll
        for (int k = 0; k < V; k += 2) {
            const float4 up = *reinterpret_cast<const float4*>(a.u + 2 * (i0 + k));
            pu[k] = up.x; pv[k] = up.y; pu[k + 1] = up.z; pv[k + 1] = up.w;
        }
    }
    const float tauI = r < 0 ? a.tauI : (float)a.rects[r].tauI;
    const float tau = r < 0 ? a.tau : (float)a.rects[r].tau;
    float tIs[NSS];
#pragma unroll
    for (int s = 0; s < NS; s++) tIs[s] = r < 0 ? a.sp[s].tauI : (float)a.rects[r].tauIs[s];
    const uint32_t it = iter_of(a.it, a.st);

    LV of[9], og[NSS][5];
    float uo[2 * V];
#pragma unroll
    for (int k = 0; k < V; k++) {
        const int xk = xl + k;
        const long i = i0 + k;
        const uint32_t dom = tg[k] & DOM_MASK;
        float d[9];
#pragma unroll
        for (int q = 0; q < 9; q++) d[q] = f[q].pulled(shift_of(q), k, je);
        float gg[NSS][5];
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++) gg[s][q] = g[s][q].pulled(shift_of(q), k, je);
        if (!LV::C::dev) {
#pragma unroll
            for (int q = 0; q < 9; q++) d[q] -= wq(q);
        }
        float m = 0.0f;
        if (a.msite) { m = (float)a.msite[i]; if (k < nvalid) a.msite[i] = 0.0; }
        else if (a.mtab) m = (dom < (uint32_t)a.ntab) ? (float)a.mtab[dom] : 0.0f;
        if (NS > 0 && (tg[k] >> BC_SHIFT)) {
            // no-flux cut links: incoming population from the first listed
            // species' pre-stream state, swapped into every listed species
            const int inv[5] = {DT, DW, DS, DE, DN};
            for (int o = 1; o <= 4; o++) {
                if (!((tg[k] >> (BC_SHIFT + o - 1)) & 1u)) continue;
                const int dd = inv[o];
                const long e = (long)(o - 1) * n + i;
                const int desc = a.bc_desc[e];
                const int b0 = a.bcd_off[desc], b1 = a.bcd_off[desc + 1];
                if (b1 <= b0) continue;
                const long s1 = (long)(5 * a.bcd_sp[b0]) * n + i;
                const float Cp = LV::dec(LV::ld1(gsb + s1)) + LV::dec(LV::ld1(gsb + s1 + n)) +
                                 LV::dec(LV::ld1(gsb + s1 + 2 * n)) + LV::dec(LV::ld1(gsb + s1 + 3 * n)) +
                                 LV::dec(LV::ld1(gsb + s1 + 4 * n));
                const float c6 = Cp / 6.0f;
                const float bvx = (float)a.bc_v[2 * e], bvy = (float)a.bc_v[2 * e + 1];
                float feq, eq;
                if (dd == DE) { feq = c6 * (1.0f + 3.0f * bvx); eq = c6 * (1.0f + 3.0f * pu[k]); }
                else if (dd == DN) { feq = c6 * (1.0f + 3.0f * bvy); eq = c6 * (1.0f + 3.0f * pv[k]); }
                else if (dd == DW) { feq = c6 * (1.0f - 3.0f * bvx); eq = c6 * (1.0f - 3.0f * pu[k]); }
                else { feq = c6 * (1.0f - 3.0f * bvy); eq = c6 * (1.0f - 3.0f * pv[k]); }
                float carry = feq + (LV::dec(LV::ld1(gsb + s1 + (long)dd * n)) - eq);
                for (int kk = b0; kk < b1; kk++) {
                    const int sk = a.bcd_sp[kk];
#pragma unroll
                    for (int s = 0; s < NS; s++)
                        if (s == sk) { const float t = gg[s][dd]; gg[s][dd] = carry; carry = t; }
                }
            }
        }
        float uxn, uyn;
        collide_dev_f32<FM>(d, m, Fx[k], Fy[k], tauI, tau, pu[k], pv[k], uxn, uyn);
        if (NS > 0) {
            float cj[NSS], cseq[NSS];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                cj[s] = gg[s][0] + gg[s][1] + gg[s][2] + gg[s][3] + gg[s][4];
                cseq[s] = cj[s];
            }
#pragma unroll
            for (int s = 0; s < NS; s++) {
                species_collide_f32<NS>(gg[s], cj[s], uxn, uyn, a.sp[s], tIs[s], dom, it, cj, cseq);
                // a later species sees this one's concentration after its collide
                cseq[s] = gg[s][0] + gg[s][1] + gg[s][2] + gg[s][3] + gg[s][4];
            }
        }
        if (a.box_outlet && (xk == fx0 || y == fy0)) {
            const int passes = (xk == fx0) + (y == fy0);
            for (int p = 0; p < passes; p++) {
                float sd = 0.0f;
#pragma unroll
                for (int q = 0; q < 9; q++) sd += d[q];
                const float sf = 1.0f / (1.0f + sd);
#pragma unroll
                for (int q = 0; q < 9; q++) d[q] = (d[q] - wq(q) * sd) * sf;   // (w + d) / (1 + sd) - w
#pragma unroll
                for (int s = 0; s < NS; s++)
#pragma unroll
                    for (int q = 0; q < 5; q++) gg[s][q] *= sf;
            }
        }
#pragma unroll
        for (int q = 0; q < 9; q++) of[q].put(k, LV::enc(LV::C::dev ? d[q] : d[q] + wq(q)));
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++) og[s][q].put(k, LV::enc(gg[s][q]));
        uo[2 * k] = uxn;
        uo[2 * k + 1] = uyn;
    }
    if (nvalid <= 0) return;
    T* fd = static_cast<T*>(a.fd);
    T* gd = static_cast<T*>(a.gd);
    if (nvalid == V) {
#pragma unroll
        for (int q = 0; q < 9; q++) of[q].store(fd + (long)q * n + i0);
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++) og[s][q].store(gd + (long)(5 * s + q) * n + i0);
#pragma unroll
        for (int k = 0; k < V; k += 2)
            *reinterpret_cast<float4*>(a.u + 2 * (i0 + k)) = make_float4(uo[2 * k], uo[2 * k + 1], uo[2 * k + 2],
                                                                          uo[2 * k + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < 9; q++) of[q].store_first(fd + (long)q * n + i0, nvalid);
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++) og[s][q].store_first(gd + (long)(5 * s + q) * n + i0, nvalid);
#pragma unroll
        for (int k = 0; k < V; k++) {
            if (k >= nvalid) break;
            *reinterpret_cast<float2*>(a.u + 2 * (i0 + k)) = make_float2(uo[2 * k], uo[2 * k + 1]);
        }
    }
}

// FP64 natural-order fields <-> storage (current buffers); field 0 = fluid.
// src/dst planes q0 .. q0 + count/n - 1 of nq in the storage blocks.  The
// deviations are taken from the exact weights, so that the rest state is
// stored as zeros and a stored state reads back with density 1 + sum(d).
__device__ __forceinline__ double wq_exact(int q) { return q == 0 ? 4.0 / 9.0 : (q < 5 ? 1.0 / 9.0 : 1.0 / 36.0); }
template <int P>
__global__ void k_pack(const double* __restrict__ src, void* dst, long count, int fluid, long n, int q0, int nq) {
    using C = Codec<P>;
    typename C::T* out = static_cast<typename C::T*>(dst);
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (long)gridDim.x * blockDim.x) {
        const int q = (int)(k / n);
        float v = (float)src[k];
        if (fluid && C::dev) v = (float)(src[k] - wq_exact(q));
        out[pidx<P>(q0 + q, k - (long)q * n, n, nq)] = C::enc(v);
    }
}
template <int P>
__global__ void k_unpack(const void* src, double* __restrict__ dst, long count, int fluid, long n, int q0, int nq) {
    using C = Codec<P>;
    const typename C::T* in = static_cast<const typename C::T*>(src);
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (long)gridDim.x * blockDim.x) {
        const int q = (int)(k / n);
        double v = (double)C::dec(in[pidx<P>(q0 + q, k - (long)q * n, n, nq)]);
        if (fluid && C::dev) v += wq_exact(q);
        dst[k] = v;
    }
}
__global__ void k_velocity_to_f32(const double* __restrict__ src, float* dst, long count) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (long)gridDim.x * blockDim.x)
        dst[k] = (float)src[k];
}


template <int NS, int FM, int P>
static void launch_mx_spt(lbib_ctx* c, const MixedArgs& a, int wg, int spt) {
    dim3 grid((c->nx + wg * spt - 1) / (wg * spt), c->ny);
    if (spt == 1) k_stream_collide_mx<NS, FM, P, 1><<<grid, wg, 0, c->stream>>>(a);
    else if (spt == 2) k_stream_collide_mx<NS, FM, P, 2><<<grid, wg, 0, c->stream>>>(a);
    else k_stream_collide_mx<NS, FM, P, 4><<<grid, wg, 0, c->stream>>>(a);
}

// Segment table of the vector kernel for workgroups of wg lanes with v sites
// each: every rectangle's row-major groups of v sites cut into runs of wg
// groups, rectangle-major so that the rows one workgroup pulls from sit in the
// same XCD's range of workgroups.  Not needed without an atlas.
static int ensure_segments(lbib_ctx* c, int wg, int v) {
    const int key = wg * 64 + v;
    if (c->nrect == 0 || (c->vsegs && c->vseg_len == key)) return 0;
    if (c->capturing) {
        c->err = "vector lattice kernel: segment table must be built before a step is recorded";
        return -1;
    }
    std::vector<SegDev> t;
    for (int r = 0; r < c->nrect; r++) {
        const int w = c->rect_host[4 * r + 2], h = c->rect_host[4 * r + 3];
        const long groups = (long)((w + v - 1) / v) * h;
        for (long l0 = 0; l0 < groups; l0 += wg) t.push_back({r, (int32_t)l0});
    }
    if (c->vsegs) { (void)hipFree(c->vsegs); c->vsegs = nullptr; c->nvseg = 0; c->vseg_len = 0; }
    c->alloc_gen++;
    LBIB_HIP(hipMalloc(&c->vsegs, t.size() * sizeof(SegDev)));
    LBIB_HIP(hipMemcpyAsync(c->vsegs, t.data(), t.size() * sizeof(SegDev), hipMemcpyHostToDevice, c->stream));
    LBIB_HIP(hipStreamSynchronize(c->stream));
    c->nvseg = (int)t.size();
    c->vseg_len = key;
    return 0;
}

template <int NS, int FM, int P, int V>
static void launch_vec(lbib_ctx* c, const MixedArgs& a, int wg) {
    if (c->nrect) {
        k_stream_collide_vec<NS, FM, P, V><<<c->nvseg, wg, 0, c->stream>>>(
            a, static_cast<const SegDev*>(c->vsegs), 0);
    } else {
        const int nchunk = (c->nx + wg * V - 1) / (wg * V);
        k_stream_collide_vec<NS, FM, P, V><<<nchunk * c->ny, wg, 0, c->stream>>>(a, nullptr, nchunk);
    }
}

template <int NS, int FM, int P>
static void launch_mx_shape(lbib_ctx* c, const MixedArgs& a, const LatticeShape& sh) {
    if (sh.variant == 2) {
        if constexpr (sizeof(typename Codec<P>::T) == 2) {
            if (sh.spt == 8) launch_vec<NS, FM, P, 8>(c, a, sh.wg);
            else launch_vec<NS, FM, P, 4>(c, a, sh.wg);
        } else {
            if (sh.spt == 4) launch_vec<NS, FM, P, 4>(c, a, sh.wg);
            else launch_vec<NS, FM, P, 2>(c, a, sh.wg);
        }
    } else if (sh.variant == 1) {
        dim3 grid((c->nx + 2 * sh.wg - 1) / (2 * sh.wg), c->ny);
        k_stream_collide_h16<NS, FM, P><<<grid, sh.wg, 0, c->stream>>>(a);
    } else {
        launch_mx_spt<NS, FM, P>(c, a, sh.wg, sh.spt);
    }
}

template <int FM, int P>
static void dispatch_mx(lbib_ctx* c, const MixedArgs& a, const LatticeShape& sh) {
    switch (c->ns) {
        case 0: launch_mx_shape<0, FM, P>(c, a, sh); break;
        case 1: launch_mx_shape<1, FM, P>(c, a, sh); break;
        case 2: launch_mx_shape<2, FM, P>(c, a, sh); break;
        case 3: launch_mx_shape<3, FM, P>(c, a, sh); break;
        default: launch_mx_shape<4, FM, P>(c, a, sh); break;
    }
}

template <int P>
static void dispatch_mx_fm(lbib_ctx* c, const MixedArgs& a, const LatticeShape& sh) {
    if (c->cfg.force_model == LBIB_FORCE_LUO) dispatch_mx<0, P>(c, a, sh);
    else if (c->cfg.force_model == LBIB_FORCE_GUO_AS_WRITTEN) dispatch_mx<1, P>(c, a, sh);
    else dispatch_mx<2, P>(c, a, sh);
}

static int launch_mixed_shape(lbib_ctx* c, const MixedArgs& a, const LatticeShape& sh) {
    if (c->prec == LBIB_FP16_REM) return launch_rem_shape(c, a, sh.variant, sh.wg);
    if (sh.variant == 2 && ensure_segments(c, sh.wg, sh.spt)) return -1;
    if (c->prec == LBIB_FP32) dispatch_mx_fm<LBIB_FP32>(c, a, sh);
    else if (c->prec == LBIB_FP16_DEV) dispatch_mx_fm<LBIB_FP16_DEV>(c, a, sh);
    else dispatch_mx_fm<LBIB_BF16_DEV>(c, a, sh);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_stream_collide_mixed(lbib_ctx* c, uint32_t it, bool use_ftile) {
    const MixedArgs a = make_mixed_args(c, it, use_ftile);
    if (launch_mixed_shape(c, a, lattice_shape(c))) return -1;
    c->cur = 1 - c->cur;
    c->mtab_live = false;
    c->msite_live = false;
    return 0;
}

static int grid_of(long count) { return (int)std::min<long>((count + 255) / 256, 8192); }

int launch_pack(lbib_ctx* c, const double* src_dev, int field, int nq) {
    if (c->prec == LBIB_FP16_REM) return launch_rem_pack(c, src_dev, field);
    void* dst = field == 0 ? c->f[c->cur] : c->g[c->cur];
    const int q0 = field == 0 ? 0 : 5 * (field - 1);
    const int nqt = field == 0 ? 9 : 5 * std::max(c->ns, 1);
    const long count = (long)nq * c->n;
    const int fl = field == 0 ? 1 : 0;
    if (c->prec == LBIB_FP32)
        k_pack<LBIB_FP32><<<grid_of(count), 256, 0, c->stream>>>(src_dev, dst, count, fl, c->n, q0, nqt);
    else if (c->prec == LBIB_FP16_DEV)
        k_pack<LBIB_FP16_DEV><<<grid_of(count), 256, 0, c->stream>>>(src_dev, dst, count, fl, c->n, q0, nqt);
    else k_pack<LBIB_BF16_DEV><<<grid_of(count), 256, 0, c->stream>>>(src_dev, dst, count, fl, c->n, q0, nqt);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_unpack(lbib_ctx* c, double* dst_dev, int field, int nq) {
    if (c->prec == LBIB_FP16_REM) return launch_rem_unpack(c, dst_dev, field);
    const void* src = field == 0 ? c->f[c->cur] : c->g[c->cur];
    const int q0 = field == 0 ? 0 : 5 * (field - 1);
    const int nqt = field == 0 ? 9 : 5 * std::max(c->ns, 1);
    const long count = (long)nq * c->n;
    const int fl = field == 0 ? 1 : 0;
    if (c->prec == LBIB_FP32)
        k_unpack<LBIB_FP32><<<grid_of(count), 256, 0, c->stream>>>(src, dst_dev, count, fl, c->n, q0, nqt);
    else if (c->prec == LBIB_FP16_DEV)
        k_unpack<LBIB_FP16_DEV><<<grid_of(count), 256, 0, c->stream>>>(src, dst_dev, count, fl, c->n, q0, nqt);
    else k_unpack<LBIB_BF16_DEV><<<grid_of(count), 256, 0, c->stream>>>(src, dst_dev, count, fl, c->n, q0, nqt);
    LBIB_HIP(hipGetLastError());
    return 0;
}

// Initialisation: the FP64 initial state (FluidSolver::initSolver /
// initWithVelocity, species initSolver) into scratch, then packed.
int launch_init_mixed(lbib_ctx* c, const double* u0_dev) {
    double *f = nullptr, *g = nullptr, *u = nullptr;
    const size_t nf = 9 * (size_t)c->n, ng = (size_t)std::max(c->ns, 1) * 5 * c->n;
    if (hipMalloc((void**)&f, nf * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&g, ng * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&u, 2 * (size_t)c->n * sizeof(double)) != hipSuccess) {
        if (f) (void)hipFree(f);
        if (g) (void)hipFree(g);
        if (u) (void)hipFree(u);
        c->err = "mixed-precision init: scratch allocation failed";
        return -1;
    }
    int r = launch_init_f64_into(c, u0_dev, f, g, u);
    if (!r) r = launch_pack(c, f, 0, 9);
    for (int s = 0; !r && s < c->ns; s++) r = launch_pack(c, g + (size_t)s * 5 * c->n, 1 + s, 5);
    if (!r) {
        k_velocity_to_f32<<<grid_of(2 * c->n), 256, 0, c->stream>>>(u, reinterpret_cast<float*>(c->u), 2 * c->n);
        if (hipGetLastError() != hipSuccess) { c->err = "mixed-precision init: launch failed"; r = -1; }
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(f);
    (void)hipFree(g);
    (void)hipFree(u);
    return r;
}

int autotune_mixed(lbib_ctx* c, int reps, std::vector<TunedShape>& out) {
    TuneScratch sc;
    if (sc.init(c, 2 * (size_t)c->n * sizeof(float))) return -1;
    // the active-tile flags are cleared after every step: the timed launches
    // read the force field at every site (the most a step reads)
    MixedArgs a = make_mixed_args(c, 0, false);
    a.u = (float*)sc.u;
    a.msite = nullptr;
    a.mtab = nullptr;
    a.st = sc.st;
    std::vector<LatticeShape> cand;
    for (int wg : {64, 128, 256})
        for (int spt : {1, 2, 4})
            if (spt == 1 || c->prec != LBIB_FP16_REM) cand.push_back({0, wg, spt});
    if (pairs_apply(c))
        for (int wg : {64, 128, 256}) cand.push_back({1, wg, 1});
    for (int v : {2, 4, 8})
        if (vec_apply(c, v))
            for (int wg : {64, 128, 256}) cand.push_back({2, wg, v});
    if (c->prec == LBIB_FP16_REM)
        for (int wg : {64, 128, 256}) cand.push_back({2, wg, 1});
    for (const LatticeShape& sh : cand) {
        double ms = 0.0;
        if (sc.time(c, reps, [&] { return launch_mixed_shape(c, a, sh); }, ms)) return -1;
        out.push_back({sh, ms});
    }
    return 0;
}

}  // namespace lbib
