// lbib_device_math.h -- per-site arithmetic shared by the FP64 kernels.
//
// Every expression keeps the operand order of the reference so that, compiled
// with -ffp-contract=off, the FP64 kernels round exactly like the reference's
// CPU code built without FMA contraction.
#pragma once
#include <hip/hip_runtime.h>
#include "lbib_internal.h"

namespace lbib {

__device__ constexpr double kW0 = 4.0 / 9.0;    // FluidSolver.cpp:39-47
__device__ constexpr double kW1 = 1.0 / 9.0;
__device__ constexpr double kW2 = 1.0 / 36.0;
__device__ constexpr double kHalfPi = M_PI / 2.0;     // GeometryNode.cpp:47-48
__device__ constexpr double kDoublePi = 2.0 * M_PI;
__device__ constexpr double kEPS = 10E-10;            // Constants.hpp:34
__device__ constexpr double kMAXLENGTH = 0.5;         // Constants.hpp:49

// Population storage formats: FP64 (the parity path), and FP32, FP16 / BF16
// (fluid populations as deviations f_i - w_i; species populations as they
// are).  dec/enc convert one stored value to and from the arithmetic type.
template <int P>
struct Codec;
template <>
struct Codec<LBIB_FP64> {
    using T = double;
    static constexpr bool dev = false;
    __device__ static double dec(T v) { return v; }
    __device__ static T enc(double x) { return x; }
};
template <>
struct Codec<LBIB_FP32> {
    // fluid populations as deviations f_i - w_i: FP32 spacing at |f - w|
    // (about 1e-9 here) instead of at w_i (up to 3e-8)
    using T = float;
    static constexpr bool dev = true;
    __device__ static float dec(T v) { return v; }
    __device__ static T enc(float x) { return x; }
};
template <>
struct Codec<LBIB_FP16_DEV> {
    using T = uint16_t;
    static constexpr bool dev = true;
    __device__ static float dec(T v) { return (float)__builtin_bit_cast(_Float16, v); }
    __device__ static T enc(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }   // round to nearest even
};
template <>
struct Codec<LBIB_BF16_DEV> {
    using T = uint16_t;
    static constexpr bool dev = true;
    __device__ static float dec(T v) { return __uint_as_float((uint32_t)v << 16); }
    __device__ static T enc(float x) {   // round to nearest even, NaN kept quiet
        uint32_t u = __float_as_uint(x);
        if ((u & 0x7fffffffu) > 0x7f800000u) return (T)((u >> 16) | 0x40u);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (T)(u >> 16);
    }
};

// Storage index of population plane q (of nq planes) at site i: planar,
// q * n + i, for every storage format.  Every kernel addresses populations
// through this one function.  A blocked layout ([i / 64][q][i % 64]) was
// measured here: a bare 14-plane 16-bit pull moved 5.8 TB/s blocked against
// 4.6-5.2 TB/s planar, but the full lattice kernels did not gain (FP16 paired
// kernel 70 -> 83 us on 2048^2), so storage stays planar; allocations keep
// whole 64-site blocks either way.
template <int P>
__device__ __forceinline__ long pidx(int q, long i, long n, int nq) {
    (void)nq;
    return (long)q * n + i;
}

// LBIB_FP16_REM: FP32 moments plus FP16 remainders.
//   fluid block:   float M[3][n] = (rho - 1, jx, jy) of the stored populations,
//                  then uint16 R[9][n] = FP16 of (f_q - w_q - deq_q(M)) * REM_SCALE
//                  with deq_q = w_q (rho - 1) + w_q rho (3 e.u + 9/2 (e.u)^2 - 3/2 u^2)
//   species block: float C[ns][n] (sum of the five populations), then uint16
//                  R[ns][5][n] = FP16 of (g_q - t_q C) * REM_SCALE, t = (1/3, 1/6 x 4)
// The moments carry FP32 precision; a remainder is small (the
// non-equilibrium part), so its FP16 rounding is small in absolute terms.
// Decoding evaluates the same FP32 expressions as encoding.  REM_SCALE lifts
// remainders of order 1e-8 out of the FP16 subnormal range; a remainder
// beyond 15.9 overflows to infinity (populations are O(1)).
constexpr float REM_SCALE = 4096.0f;
constexpr float REM_INV = 1.0f / 4096.0f;
__device__ __forceinline__ float rem_w(int q) { return q == 0 ? 4.0f / 9.0f : (q < 5 ? 1.0f / 9.0f : 1.0f / 36.0f); }
__device__ __forceinline__ float rem_t(int q) { return q == 0 ? 1.0f / 3.0f : 1.0f / 6.0f; }
__device__ __forceinline__ int rem_ex(int q) { return (q == 1 || q == 5 || q == 8) ? 1 : ((q == 3 || q == 6 || q == 7) ? -1 : 0); }
__device__ __forceinline__ int rem_ey(int q) { return (q == 2 || q == 5 || q == 6) ? 1 : ((q == 4 || q == 7 || q == 8) ? -1 : 0); }
// equilibrium deviation deq_q of the moments (rho - 1, jx, jy)
__device__ __forceinline__ float rem_deq(int q, float rm1, float jx, float jy) {
    const float rho = 1.0f + rm1;
    const float rhoI = 1.0f / rho;
    const float ux = jx * rhoI, uy = jy * rhoI;
    const float u2 = 1.5f * (ux * ux + uy * uy);
    const float eu = (float)rem_ex(q) * ux + (float)rem_ey(q) * uy;
    const float w = rem_w(q);
    return w * rm1 + w * rho * (3.0f * eu + 4.5f * eu * eu - u2);
}
__device__ __forceinline__ uint16_t rem_enc(float r) { return __builtin_bit_cast(uint16_t, (_Float16)(r * REM_SCALE)); }
__device__ __forceinline__ float rem_dec(uint16_t v) { return (float)__builtin_bit_cast(_Float16, v) * REM_INV; }
// moments of fluid deviations d[9] (the sums the encoder stores)
__device__ __forceinline__ void rem_moments(const float d[9], float& rm1, float& jx, float& jy) {
    rm1 = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7])) + d[8];
    jx = (d[1] + d[5] + d[8]) - (d[3] + d[6] + d[7]);
    jy = (d[2] + d[5] + d[6]) - (d[4] + d[7] + d[8]);
}
__device__ __forceinline__ float rem_csum(const float g[5]) { return (g[0] + g[1]) + (g[2] + g[3]) + g[4]; }

// Species populations of one storage block of nq = 5 * max(ns, 1) planes, for
// code that handles one site at a time (re-initialisation, per-domain sums):
// get returns a population in FP64, put5 stores the five of one species.
template <int P>
struct GStore {
    __device__ static double get(const void* g, int s, int q, long i, long n, int nq) {
        return (double)Codec<P>::dec(static_cast<const typename Codec<P>::T*>(g)[pidx<P>(5 * s + q, i, n, nq)]);
    }
    __device__ static void put5(void* g, int s, long i, long n, int nq, const double d[5]) {
        typename Codec<P>::T* gb = static_cast<typename Codec<P>::T*>(g);
        for (int q = 0; q < 5; q++) gb[pidx<P>(5 * s + q, i, n, nq)] = Codec<P>::enc(d[q]);
    }
};
template <>
struct GStore<LBIB_FP16_REM> {
    __device__ static double get(const void* g, int s, int q, long i, long n, int nq) {
        const float* C = static_cast<const float*>(g);
        const uint16_t* R = reinterpret_cast<const uint16_t*>(C + (long)(nq / 5) * n);
        return (double)(rem_t(q) * C[(long)s * n + i] + rem_dec(R[(long)(5 * s + q) * n + i]));
    }
    __device__ static void put5(void* g, int s, long i, long n, int nq, const double d[5]) {
        float* C = static_cast<float*>(g);
        uint16_t* R = reinterpret_cast<uint16_t*>(C + (long)(nq / 5) * n);
        float gf[5];
        for (int q = 0; q < 5; q++) gf[q] = (float)d[q];
        const float c = rem_csum(gf);
        C[(long)s * n + i] = c;
        for (int q = 0; q < 5; q++) R[(long)(5 * s + q) * n + i] = rem_enc(gf[q] - rem_t(q) * c);
    }
};

// GeometryNode.cpp:55-81 (parabolic sine approximation, not cos)
__device__ __forceinline__ double fast_cos(const double x) {
    double c;
    double t = x + kHalfPi;
    if (t > M_PI) t -= kDoublePi;
    if (t < .00) {
        c = 1.27323954 * t + 0.405284735 * t * t;
    } else {
        c = 1.27323954 * t - 0.405284735 * t * t;
    }
    if (c < 0.0) c = 0.225 * (c * -c - c) + c;
    else         c = 0.225 * (c * c - c) + c;
    return c;
}
// GeometryNode.cpp:89-94; dx = x_site - X_marker
__device__ __forceinline__ double ib_delta(const double dx, const double dy) {
    return ((1.0 + fast_cos(kHalfPi * dx)) * (1.0 + fast_cos(kHalfPi * dy))) / 16.0;
}
// The two factors of ib_delta for the four stencil offsets along one axis:
// f[q] = 1 + fastCos(pi/2 * ((s0 + q) - X)), so that
// ib_delta((s0x + qx) - X, (s0y + qy) - Y) == (fx[qx] * fy[qy]) / 16 bit for bit.
__device__ __forceinline__ void ib_factors(const int s0, const double X, double f[4]) {
#pragma unroll
    for (int q = 0; q < 4; q++) f[q] = 1.0 + fast_cos(kHalfPi * ((double)(s0 + q) - X));
}

// FluidSolver::collide (FluidSolver.cpp:103-183) on natural-order pulled
// populations d[9]; Luo1993ForceModel (Luo1993ForceModel.cpp:29-70) or
// GuoZhengShi2002ForceModel (GuoZhengShi2002ForceModel.cpp:29-78).
// gx,gy: stored velocity of the previous collide (Guo only).  Writes the
// post-collision velocity to ux,uy.
template <int FM>
__device__ __forceinline__ void fluid_collide(double d[9], const double m, const double Fx,
                                              const double Fy, const double tauI, const double tau,
                                              const double gx, const double gy, double& uxo,
                                              double& uyo) {
    double rho0 = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) rho0 += d[q];
    double rho = rho0;
    if (isfinite(m)) rho += m;
    double ux, uy;
    {
        const double rhoI = 1.0 / rho;
        if (FM == 0) {
            ux = (d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW])) * rhoI;
            uy = (d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE])) * rhoI;
        } else {
            ux = (d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW]) + 0.5 * Fx) * rhoI;
            uy = (d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE]) + 0.5 * Fy) * rhoI;
        }
    }
    const double ux2 = ux * ux, uy2 = uy * uy;
    const double lastTerm = 3.0 / 2.0 * (ux2 + uy2);
    const double W3 = 9.0 / 2.0;
    const double W1rho = kW1 * rho, W2rho = kW2 * rho;
    double fc[9];
    if (FM == 0) {
        const double W1rho3 = rho0 * 3.0 / 9.0, W2rho3 = rho0 * 3.0 / 36.0;
        fc[DT] = 0.0;
        fc[DE] = W1rho3 * Fx;
        fc[DN] = W1rho3 * Fy;
        fc[DW] = -W1rho3 * Fx;
        fc[DS] = -W1rho3 * Fy;
        fc[DNE] = W2rho3 * (Fx + Fy);
        fc[DNW] = W2rho3 * (-Fx + Fy);
        fc[DSW] = W2rho3 * (-Fx - Fy);
        fc[DSE] = W2rho3 * (Fx - Fy);
    } else {
        const double G0 = (1.0 - 0.5 / tau) * 4.0 / 9.0;
        const double G1 = (1.0 - 0.5 / tau) / 9.0;
        const double G2 = (1.0 - 0.5 / tau) / 36.0;
        fc[DT] = G0 * (-3.0 * gx * Fx - 3.0 * gy * Fy);
        fc[DE] = G1 * ((6.0 * gx + 3.0) * Fx - 3.0 * gy * Fy);
        fc[DN] = G1 * (-3.0 * gx * Fx + (6.0 * gy + 3.0) * Fy);
        fc[DW] = G1 * ((6.0 * gx - 3.0) * Fx - 3.0 * gy * Fy);
        fc[DS] = G1 * (-3.0 * gx * Fx + (6.0 * gy - 3.0) * Fy);
        if (FM == 1)
            fc[DNE] = G2 * ((6.0 * gx + 9.0 * gy + 3.0) * Fx + (9.0 * gx - 3.0 * gy + 3.0) * Fy);
        else
            fc[DNE] = G2 * ((6.0 * gx + 9.0 * gy + 3.0) * Fx + (9.0 * gx + 6.0 * gy + 3.0) * Fy);
        fc[DNW] = G2 * ((6.0 * gx - 9.0 * gy - 3.0) * Fx + (-9.0 * gx + 6.0 * gy + 3.0) * Fy);
        fc[DSW] = G2 * ((6.0 * gx + 9.0 * gy - 3.0) * Fx + (9.0 * gx + 6.0 * gy - 3.0) * Fy);
        fc[DSE] = G2 * ((6.0 * gx - 9.0 * gy + 3.0) * Fx + (-9.0 * gx + 6.0 * gy - 3.0) * Fy);
    }
    double tmp[9];
    tmp[DT] = kW0 * rho * (1.0 - lastTerm) * tauI + fc[DT];
    tmp[DE] = W1rho * (1.0 + 3.0 * ux + W3 * ux2 - lastTerm) * tauI + fc[DE];
    tmp[DNE] = W2rho * (1.0 + 3.0 * (ux + uy) + W3 * (ux + uy) * (ux + uy) - lastTerm) * tauI + fc[DNE];
    tmp[DN] = W1rho * (1.0 + 3.0 * uy + W3 * uy2 - lastTerm) * tauI + fc[DN];
    tmp[DNW] = W2rho * (1.0 + 3.0 * (-ux + uy) + W3 * (-ux + uy) * (-ux + uy) - lastTerm) * tauI + fc[DNW];
    tmp[DW] = W1rho * (1.0 + 3.0 * (-ux) + W3 * ux2 - lastTerm) * tauI + fc[DW];
    tmp[DSW] = W2rho * (1.0 + 3.0 * (-ux - uy) + W3 * (-ux - uy) * (-ux - uy) - lastTerm) * tauI + fc[DSW];
    tmp[DS] = W1rho * (1.0 + 3.0 * (-uy) + W3 * uy2 - lastTerm) * tauI + fc[DS];
    tmp[DSE] = W2rho * (1.0 + 3.0 * (ux - uy) + W3 * (ux - uy) * (ux - uy) - lastTerm) * tauI + fc[DSE];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const double td = d[q];
        d[q] = td - td * tauI + tmp[q];
    }
    const double rhoI = 1.0 / rho;
    if (FM == 0) {
        uxo = (d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW])) * rhoI;
        uyo = (d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE])) * rhoI;
    } else {
        uxo = (d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW]) + 0.5 * Fx) * rhoI;
        uyo = (d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE]) + 0.5 * Fy) * rhoI;
    }
}

// Reaction term (…_SIGNAL.cpp:92-100 for SIGNAL; the linear spec is
// evaluated on the pre-collision concentrations cj[] of every species; the
// Schnakenberg pair reads its partner's concentration from cseq[], which holds
// each species' C as the reference's getC() sees it at that moment: after the
// partner's collide (slot-order sum) if it comes earlier in the species
// order, before it otherwise).
template <int NS>
__device__ __forceinline__ double species_reaction(const SpeciesDev& p, const uint32_t dom,
                                                   const uint32_t it, const double C,
                                                   const double* cj, const double* cseq) {
    if (p.kind == LBIB_SPECIES_SCHNAKENBERG_R || p.kind == LBIB_SPECIES_SCHNAKENBERG_L) {
        if (dom == 0u) return 0.0;
        const double Cp = cseq[p.prod_domain];
        if (p.kind == LBIB_SPECIES_SCHNAKENBERG_R) return p.production * (p.source_inside - C + C * C * Cp);
        return p.production * (p.source_inside - Cp * Cp * C);
    }
    if (p.kind == LBIB_SPECIES_SIGNAL) {
        const bool prod = (p.prod_domain < 0) ? (dom != 0u) : (dom == (uint32_t)p.prod_domain);
        if (prod && (it < p.switchoff)) return p.production;
        return -p.decay * C;
    }
    if (p.kind == LBIB_SPECIES_LINEAR) {
        double r = 0.0;
#pragma unroll
        for (int j = 0; j < NS; j++) r += p.krow[j] * cj[j];
        if (dom != 0u) r += p.source_inside;
        return r;
    }
    return 0.0;
}

// CDESolverD2Q5::collide (CDESolverD2Q5.cpp:75-107) / SIGNAL (…_SIGNAL.cpp:102-135)
// on natural-order d[5]; C must be d[0]+…+d[4] in index order.  tauI is the
// species' 1/tau in the site's domain.
template <int NS>
__device__ __forceinline__ void species_collide(double d[5], const double C, const double ux,
                                                const double uy, const SpeciesDev& p, const double tauI,
                                                const uint32_t dom, const uint32_t it,
                                                const double* cj, const double* cseq) {
    const double w0 = C / 3.0, w1 = C / 6.0;
    double tmp[5];
    tmp[DT] = w0;
    tmp[DE] = w1 * (1.0 + ux * 3.0);
    tmp[DN] = w1 * (1.0 + uy * 3.0);
    tmp[DW] = w1 * (1.0 + (-ux) * 3.0);
    tmp[DS] = w1 * (1.0 + (-uy) * 3.0);
    if (p.kind == LBIB_SPECIES_PLAIN) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const double td = d[q];
            d[q] = td - td * tauI + tmp[q] * tauI;
        }
    } else {
        const double R = species_reaction<NS>(p, dom, it, C, cj, cseq);
        const bool sch = p.kind == LBIB_SPECIES_SCHNAKENBERG_R || p.kind == LBIB_SPECIES_SCHNAKENBERG_L;
        const double R0 = sch ? p.decay * R / 3.0 : 1.0 * R / 3.0;
        const double R1 = sch ? p.decay * R / 6.0 : 1.0 * R / 6.0;
        d[DT] = d[DT] - d[DT] * tauI + tmp[DT] * tauI + R0;
#pragma unroll
        for (int q = 1; q < 5; q++) d[q] = d[q] - d[q] * tauI + tmp[q] * tauI + R1;
    }
}

// Sum of the 9 stored values in the reference's swapped slot order
// (FluidSolver::getRho on the locally swapped state): T,W,S,E,N,SW,SE,NE,NW.
__device__ __forceinline__ double rho_swapped(const double d[9]) {
    double r = 0.0;
    r += d[DT]; r += d[DW]; r += d[DS]; r += d[DE]; r += d[DN];
    r += d[DSW]; r += d[DSE]; r += d[DNE]; r += d[DNW];
    return r;
}
// CDESolverD2Q5::getC on the swapped state: T,W,S,E,N.
__device__ __forceinline__ double c_swapped(const double d0, const double dE, const double dN,
                                            const double dW, const double dS) {
    double r = 0.0;
    r += d0; r += dW; r += dS; r += dE; r += dN;
    return r;
}

// Wave-aggregated counter allocation: every lane of the wavefront must call
// it (v = 0 for idle lanes); returns the lane's base in *ctr.
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Renumbering them so that every XCD works on a contiguous range keeps
// neighbouring data (halo rows, stencil sites of adjacent markers) in one L2.
// A bijection of [0, total) for any total.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t lid, uint32_t total) {
    constexpr uint32_t NXCD = 8;
    const uint32_t q = total / NXCD, r = total % NXCD;
    const uint32_t x = lid % NXCD, local = lid / NXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + local;
}

// Tile rows dealt to the 8 XCDs in turn (row r to XCD r % 8), consecutive
// workgroups of one XCD walking along a row, for a grid of
// 8 * ceil(nty / 8) * ntx workgroups; -1 for the workgroups past the last row.
__device__ __forceinline__ int row_interleaved_tile(uint32_t lid, int ntx, int nty) {
    const int x = (int)(lid % 8u), local = (int)(lid / 8u);
    const int row = x + 8 * (local / ntx), col = local % ntx;
    return row < nty ? row * ntx + col : -1;
}

// Inclusive prefix sum across a 64-lane wavefront.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive prefix sum across the workgroup (blockDim.x a multiple of 64, at
// most 1024); wsum: 16 words of LDS.  Returns the workgroup total.  Every
// thread must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* excl, uint32_t* wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint32_t t = (lane < nw) ? wsum[lane] : 0u;
        t = wave_incl_scan(t);
        if (lane < nw) wsum[lane] = t;
    }
    __syncthreads();
    *excl = (w ? wsum[w - 1] : 0u) + inc - v;
    const uint32_t total = wsum[nw - 1];
    __syncthreads();
    return total;
}

__device__ __forceinline__ uint32_t wave_alloc(uint32_t* ctr, uint32_t v) {
    const int lane = threadIdx.x & 63;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    uint32_t base = 0;
    if (lane == 63 && total) base = atomicAdd(ctr, total);
    base = __shfl(base, 63, 64);
    return base + incl - v;
}

// Block-aggregated counter allocation (blockDim.x a multiple of 64, at most
// 1024): one atomic per workgroup.  Atomics on one address serialise (about
// 10 ns each on MI355X), so counters shared by the whole grid are advanced at
// most once per workgroup.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_alloc(uint32_t* ctr, uint32_t v) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t bbase;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int k = 0; k < nw; k++) total += wsum[k];
        bbase = total ? atomicAdd(ctr, total) : 0u;
    }
    __syncthreads();
    uint32_t off = bbase;
    for (int k = 0; k < w; k++) off += wsum[k];
    const uint32_t r = off + incl - v;
    __syncthreads();
    return r;
}

}  // namespace lbib
