// lbib_adjoint.hip -- discrete adjoint of the coupled D2Q9 + D2Q5 + IB step.
//
// The differentiable model is the coupled lattice step of SimulationRunner
// (SimulationRunner.cpp:72-100) restricted to its continuous part, on one
// periodic nx x ny lattice:
//   marker forces   F_k = kappa * sum over both ring neighbours (X_j - X_k)/|X_j - X_k|
//                   (ForceType6, ForceStructs.hpp:455-514, tension renewed on every
//                   connection as tutorial_01_BioSolverMembraneTension.cpp:42-77)
//   spread          F(x) = sum_k delta(x - X_k) F_k            (GeometryNode.cpp:228-235)
//   fluid           pull streaming + BGK with Luo forcing     (FluidSolver.cpp:103-183,
//                   Luo1993ForceModel.cpp:29-70), no mass source
//   junctions       F_k += k_j (|d| - l0) d / |d|, d = X_m - X_k, over a fixed list
//                   of marker pairs of different cells (ForceType7,
//                   ForceStructs.hpp:521-588, as the cell-junction BioSolver sets it)
//   species         1-4 species; pull streaming + D2Q5 BGK on the fresh velocity
//                   with the reaction R_s = sum_j K_sj C_j, plus s[c] for species 0
//                   inside cell c, and -decay_s * C_s outside every cell
//                   (tutorial_01_CDESolverD2Q5_SIGNAL.cpp:92-135 with one rate per
//                   cell; the linear reaction matrix of the multi-species runs)
//   interpolation   U_k = sum_x delta(x - X_k) u(x), X_k += U_k  (GeometryNode.cpp:216-254)
// with the fastCos delta of GeometryNode.cpp:55-94.  The cell map (which
// cell's rate applies at a site) is held at its initial value, and cut links,
// re-initialisation and remeshing are not part of the model (they are
// discrete).  The loss is L = sum_x (C_T(x) - C*(x))^2 on the final
// concentrations of all species; the parameters are tau_c per species,
// kappa, k_j, the reaction matrix K and the per-cell rates s[c].
//
// A batch of independent problems may share the lattice as periodic
// rectangles (shared parameters, summed loss).
//
// Checkpointing: either states every K steps, the reverse sweep restoring a
// checkpoint, recomputing its segment with every intermediate state stored
// and running the adjoint kernels backwards over it; or binomial
// checkpointing (revolve, Griewank & Walther's Algorithm 799) in a fixed
// number of snapshot slots, recomputing each step at most r times for
// C(s + r, s) steps.  Populations may be stored as 16-bit deviations with FP32
// arithmetic (the forward model of the mixed-precision lattice), which halves
// every snapshot.
//
// Roofline: every kernel here is HBM bound (the per-site adjoint reads the
// forward state f, g, F, u and the adjoints of step t+1 and writes those of
// step t: about 2.5x the forward lattice kernel's bytes).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>
#include "lbib_device_math.h"

struct lbib_adj;

namespace lbib {
namespace adj {

template <typename T>
struct Cst {
    static constexpr T halfpi = (T)(M_PI / 2.0);
    static constexpr T pi = (T)M_PI;
    static constexpr T twopi = (T)(2.0 * M_PI);
    static constexpr T A = (T)1.27323954;
    static constexpr T B = (T)0.405284735;
    static constexpr T W0 = (T)(4.0 / 9.0);
    static constexpr T W1 = (T)(1.0 / 9.0);
    static constexpr T W2 = (T)(1.0 / 36.0);
};

// fastCos (GeometryNode.cpp:55-81) and its derivative.
template <typename T>
__device__ __forceinline__ void fast_cos_d(const T x, T& c, T& dc) {
    T t = x + Cst<T>::halfpi;
    if (t > Cst<T>::pi) t -= Cst<T>::twopi;
    T c1, d1;
    if (t < (T)0) {
        c1 = Cst<T>::A * t + Cst<T>::B * t * t;
        d1 = Cst<T>::A + (T)2 * Cst<T>::B * t;
    } else {
        c1 = Cst<T>::A * t - Cst<T>::B * t * t;
        d1 = Cst<T>::A - (T)2 * Cst<T>::B * t;
    }
    T d2;
    if (c1 < (T)0) {
        c = (T)0.225 * (c1 * -c1 - c1) + c1;
        d2 = (T)0.225 * ((T)-2 * c1 - (T)1) + (T)1;
    } else {
        c = (T)0.225 * (c1 * c1 - c1) + c1;
        d2 = (T)0.225 * ((T)2 * c1 - (T)1) + (T)1;
    }
    dc = d2 * d1;
}

// The 4 stencil factors h_q = 1 + fastCos(pi/2 (s0 + q - X)) along one axis
// (delta = hx * hy / 16) and their derivatives with respect to X.
template <typename T>
__device__ __forceinline__ void stencil(const T X, int& s0, T h[4], T dh[4]) {
    s0 = (int)floor((double)X) - 1;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        T c, dc;
        fast_cos_d<T>(Cst<T>::halfpi * ((T)(s0 + q) - X), c, dc);
        h[q] = (T)1 + c;
        dh[q] = -Cst<T>::halfpi * dc;
    }
}

__device__ __forceinline__ long wrap_site(int x, int y, int nx, int ny) {
    x %= nx; if (x < 0) x += nx;
    y %= ny; if (y < 0) y += ny;
    return (long)y * nx + x;
}

// A batch of independent problems: periodic rectangles of one lattice (an
// atlas).  rtile holds the rectangle of every 16x16 tile (-1: none); sites
// outside every rectangle are not updated and carry no loss.  Null rtile:
// one periodic lattice.
struct AdjFrames {
    const int32_t* rtile;
    const int4* rects;   // x0, y0, w, h
    int ntx;
};
// periodic frame of site (x, y); false when no rectangle holds the site
__device__ __forceinline__ bool adj_frame(const AdjFrames& f, int x, int y, int nx, int ny, int4& fr) {
    if (!f.rtile) {
        fr = make_int4(0, 0, nx, ny);
        return true;
    }
    const int r = f.rtile[(y >> 4) * f.ntx + (x >> 4)];
    if (r < 0) return false;
    fr = f.rects[r];
    return x - fr.x < fr.z && y - fr.y < fr.w;
}

// Stencil sites of marker k: in a batch each marker wraps inside its own
// rectangle (mrect[k]), as the lattice kernels wrap each problem's pulls, so a
// stencil near a rectangle edge never reaches the gap or another problem.
struct MarkerFrames {
    const int32_t* mrect;   // rectangle per marker, null: one periodic lattice
    const int4* rects;
};
struct MarkerWrap {
    int x0, y0, w, h, nx;
    __device__ MarkerWrap(const MarkerFrames& mf, int k, int nx_, int ny_) : nx(nx_) {
        if (mf.mrect) {
            const int4 r = mf.rects[mf.mrect[k]];
            x0 = r.x; y0 = r.y; w = r.z; h = r.w;
        } else {
            x0 = 0; y0 = 0; w = nx_; h = ny_;
        }
    }
    __device__ long site(int x, int y) const {
        int xr = (x - x0) % w; if (xr < 0) xr += w;
        int yr = (y - y0) % h; if (yr < 0) yr += h;
        return (long)(y0 + yr) * nx + (x0 + xr);
    }
};

// ---------------------------------------------------------------- forward --
// Junction springs of marker k: partners jpart[jofs[k] .. jofs[k + 1]) (a pair
// is listed at both ends, so every force is gathered, never scattered).
struct Junctions {
    const int32_t* ofs;
    const int32_t* part;
    double kj, l0;
};

template <typename T>
__global__ void k_force(int nm, const T* __restrict__ X, const int32_t* __restrict__ nxt,
                        const int32_t* __restrict__ prv, T kappa, T* __restrict__ Fm, const Junctions jn) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const T x1 = X[2 * k], y1 = X[2 * k + 1];
    T fx = 0, fy = 0;
    const int nb[2] = {nxt[k], prv[k]};
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const T x2 = X[2 * nb[s]], y2 = X[2 * nb[s] + 1];
        const T dist = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
        fx += kappa * (x2 - x1) / dist;
        fy += kappa * (y2 - y1) / dist;
    }
    if (jn.ofs) {
        const T kj = (T)jn.kj, l0 = (T)jn.l0;
        for (int p = jn.ofs[k]; p < jn.ofs[k + 1]; p++) {
            const int m = jn.part[p];
            const T dx = X[2 * m] - x1, dy = X[2 * m + 1] - y1;
            const T dist = sqrt(dx * dx + dy * dy);
            const T c = kj * (dist - l0) / dist;
            fx += c * dx;
            fy += c * dy;
        }
    }
    Fm[2 * k] = fx;
    Fm[2 * k + 1] = fy;
}

template <typename T>
__global__ void k_spread(int nm, const T* __restrict__ X, const T* __restrict__ Fm, T* Fs, int nx, int ny,
                         const MarkerFrames mf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const MarkerWrap wr(mf, k, nx, ny);
    int sx, sy;
    T hx[4], hy[4], dhx[4], dhy[4];
    stencil<T>(X[2 * k], sx, hx, dhx);
    stencil<T>(X[2 * k + 1], sy, hy, dhy);
    const T fx = Fm[2 * k], fy = Fm[2 * k + 1];
#pragma unroll
    for (int qx = 0; qx < 4; qx++)
#pragma unroll
        for (int qy = 0; qy < 4; qy++) {
            const T d = (hx[qx] * hy[qy]) / (T)16;
            const long s = wr.site(sx + qx, sy + qy);
            atomicAdd(&Fs[2 * s], d * fx);
            atomicAdd(&Fs[2 * s + 1], d * fy);
        }
}

// Population storage of the states: SP == 0 keeps them in the arithmetic
// type; LBIB_FP16_DEV / LBIB_BF16_DEV keep 16-bit deviations (fluid f - w_q,
// species as they are, their rest state being 0), decoded to the arithmetic
// type on load and rounded to nearest on store.  The reverse sweep reads the
// same stored states and takes the rounding's derivative as 1.
template <typename T, int SP>
struct PopStore {
    using R = uint16_t;
    __device__ static T wgt(int q) { return q == 0 ? Cst<T>::W0 : (q < 5 ? Cst<T>::W1 : Cst<T>::W2); }
    __device__ static T f(const void* p, long k, int q) {
        return (T)Codec<SP>::dec(static_cast<const R*>(p)[k]) + wgt(q);
    }
    __device__ static void put_f(void* p, long k, int q, T v) {
        static_cast<R*>(p)[k] = Codec<SP>::enc((float)(v - wgt(q)));
    }
    __device__ static T g(const void* p, long k) { return (T)Codec<SP>::dec(static_cast<const R*>(p)[k]); }
    __device__ static void put_g(void* p, long k, T v) { static_cast<R*>(p)[k] = Codec<SP>::enc((float)v); }
};
template <typename T>
struct PopStore<T, 0> {
    __device__ static T f(const void* p, long k, int) { return static_cast<const T*>(p)[k]; }
    __device__ static void put_f(void* p, long k, int, T v) { static_cast<T*>(p)[k] = v; }
    __device__ static T g(const void* p, long k) { return static_cast<const T*>(p)[k]; }
    __device__ static void put_g(void* p, long k, T v) { static_cast<T*>(p)[k] = v; }
};

template <typename T>
struct LatArgs {
    int nx, ny;
    long n;
    const void* fs;   // states in the population storage (PopStore)
    const void* gs;
    void* fd;
    void* gd;
    const T* Fs;
    T* u;
    const uint32_t* dom;
    const double* src;
    T tauI;
    T tauIc[LBIB_MAX_SPECIES], decay[LBIB_MAX_SPECIES];
    T Km[LBIB_MAX_SPECIES * LBIB_MAX_SPECIES];   // reaction matrix, row-major (stride LBIB_MAX_SPECIES)
    int hasK;                                     // any nonzero entry
    AdjFrames fr;
};

// pull indices of the 9 directions (natural order) at site (x, y), periodic
// in the frame fr = (x0, y0, w, h) of a lattice nx wide
__device__ __forceinline__ void pull_index(int x, int y, int nx, const int4 fr, long idx[9]) {
    const int xm = x == fr.x ? fr.x + fr.z - 1 : x - 1, xp = x == fr.x + fr.z - 1 ? fr.x : x + 1;
    const int ym = y == fr.y ? fr.y + fr.w - 1 : y - 1, yp = y == fr.y + fr.w - 1 ? fr.y : y + 1;
    const long r0 = (long)y * nx, rS = (long)ym * nx, rN = (long)yp * nx;
    idx[DT] = r0 + x;
    idx[DE] = r0 + xm;
    idx[DN] = rS + x;
    idx[DW] = r0 + xp;
    idx[DS] = rN + x;
    idx[DNE] = rS + xm;
    idx[DNW] = rS + xp;
    idx[DSW] = rN + xp;
    idx[DSE] = rN + xm;
}

// FluidSolver::collide with the Luo force (same operand order as the FP64
// kernel, m = 0) and the D2Q5 collides with the per-cell reaction and the
// reaction matrix (concentrations of every species before the collides).
template <typename T, int NS>
__device__ __forceinline__ void collide_site(T d[9], T g[NS][5], const T Fx, const T Fy, const LatArgs<T>& a,
                                             const uint32_t dm, T& uxo, T& uyo) {
    const T tauI = a.tauI;
    using Q = Cst<T>;
    T rho0 = 0;
#pragma unroll
    for (int q = 0; q < 9; q++) rho0 += d[q];
    const T rho = rho0;   // no mass source
    const T rhoI = (T)1 / rho;
    const T ux = (d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW])) * rhoI;
    const T uy = (d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE])) * rhoI;
    const T ux2 = ux * ux, uy2 = uy * uy;
    const T lastTerm = (T)3 / (T)2 * (ux2 + uy2);
    const T W3 = (T)9 / (T)2;
    const T W1rho = Q::W1 * rho, W2rho = Q::W2 * rho;
    const T W1rho3 = rho0 * (T)3 / (T)9, W2rho3 = rho0 * (T)3 / (T)36;
    T fc[9];
    fc[DT] = 0;
    fc[DE] = W1rho3 * Fx;
    fc[DN] = W1rho3 * Fy;
    fc[DW] = -W1rho3 * Fx;
    fc[DS] = -W1rho3 * Fy;
    fc[DNE] = W2rho3 * (Fx + Fy);
    fc[DNW] = W2rho3 * (-Fx + Fy);
    fc[DSW] = W2rho3 * (-Fx - Fy);
    fc[DSE] = W2rho3 * (Fx - Fy);
    T tmp[9];
    tmp[DT] = Q::W0 * rho * ((T)1 - lastTerm) * tauI + fc[DT];
    tmp[DE] = W1rho * ((T)1 + (T)3 * ux + W3 * ux2 - lastTerm) * tauI + fc[DE];
    tmp[DNE] = W2rho * ((T)1 + (T)3 * (ux + uy) + W3 * (ux + uy) * (ux + uy) - lastTerm) * tauI + fc[DNE];
    tmp[DN] = W1rho * ((T)1 + (T)3 * uy + W3 * uy2 - lastTerm) * tauI + fc[DN];
    tmp[DNW] = W2rho * ((T)1 + (T)3 * (-ux + uy) + W3 * (-ux + uy) * (-ux + uy) - lastTerm) * tauI + fc[DNW];
    tmp[DW] = W1rho * ((T)1 + (T)3 * (-ux) + W3 * ux2 - lastTerm) * tauI + fc[DW];
    tmp[DSW] = W2rho * ((T)1 + (T)3 * (-ux - uy) + W3 * (-ux - uy) * (-ux - uy) - lastTerm) * tauI + fc[DSW];
    tmp[DS] = W1rho * ((T)1 + (T)3 * (-uy) + W3 * uy2 - lastTerm) * tauI + fc[DS];
    tmp[DSE] = W2rho * ((T)1 + (T)3 * (ux - uy) + W3 * (ux - uy) * (ux - uy) - lastTerm) * tauI + fc[DSE];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const T td = d[q];
        d[q] = td - td * tauI + tmp[q];
    }
    uxo = (d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW])) * rhoI;
    uyo = (d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE])) * rhoI;
    T Cs[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        T C = 0;
#pragma unroll
        for (int q = 0; q < 5; q++) C += g[s][q];
        Cs[s] = C;
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const T C = Cs[s], tauIc = a.tauIc[s];
        const T w0 = C / (T)3, w1 = C / (T)6;
        T te[5];
        te[DT] = w0;
        te[DE] = w1 * ((T)1 + uxo * (T)3);
        te[DN] = w1 * ((T)1 + uyo * (T)3);
        te[DW] = w1 * ((T)1 + (-uxo) * (T)3);
        te[DS] = w1 * ((T)1 + (-uyo) * (T)3);
        T R = dm != 0u ? (s == 0 ? (T)a.src[dm] : (T)0) : -a.decay[s] * C;
        if (a.hasK) {
#pragma unroll
            for (int j = 0; j < NS; j++) R += a.Km[s * LBIB_MAX_SPECIES + j] * Cs[j];
        }
        const T R0 = (T)1 * R / (T)3, R1 = (T)1 * R / (T)6;
        g[s][DT] = g[s][DT] - g[s][DT] * tauIc + te[DT] * tauIc + R0;
#pragma unroll
        for (int q = 1; q < 5; q++) g[s][q] = g[s][q] - g[s][q] * tauIc + te[q] * tauIc + R1;
    }
}

template <typename T, int NS, int SP>
__global__ void k_lattice(const LatArgs<T> a) {
    using PS = PopStore<T, SP>;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int x = (int)(i % a.nx), y = (int)(i / a.nx);
    int4 fr;
    if (!adj_frame(a.fr, x, y, a.nx, a.ny, fr)) return;
    long idx[9];
    pull_index(x, y, a.nx, fr, idx);
    const long n = a.n;
    T d[9], g[NS][5];
#pragma unroll
    for (int q = 0; q < 9; q++) d[q] = PS::f(a.fs, q * n + idx[q], q);
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) g[s][q] = PS::g(a.gs, (5 * s + q) * n + idx[q]);
    T uxo, uyo;
    collide_site<T, NS>(d, g, a.Fs[2 * i], a.Fs[2 * i + 1], a, a.dom[i], uxo, uyo);
#pragma unroll
    for (int q = 0; q < 9; q++) PS::put_f(a.fd, q * n + i, q, d[q]);
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) PS::put_g(a.gd, (5 * s + q) * n + i, g[s][q]);
    a.u[2 * i] = uxo;
    a.u[2 * i + 1] = uyo;
}

template <typename T>
__global__ void k_interp_move(int nm, T* X, const T* __restrict__ u, int nx, int ny, const MarkerFrames mf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const MarkerWrap wr(mf, k, nx, ny);
    int sx, sy;
    T hx[4], hy[4], dhx[4], dhy[4];
    stencil<T>(X[2 * k], sx, hx, dhx);
    stencil<T>(X[2 * k + 1], sy, hy, dhy);
    T vx = 0, vy = 0;
#pragma unroll
    for (int qx = 0; qx < 4; qx++)
#pragma unroll
        for (int qy = 0; qy < 4; qy++) {
            const T d = (hx[qx] * hy[qy]) / (T)16;
            const long s = wr.site(sx + qx, sy + qy);
            vx += d * u[2 * s];
            vy += d * u[2 * s + 1];
        }
    X[2 * k] += vx;
    X[2 * k + 1] += vy;
}

// loss L = sum over species and sites of (C - C*)^2 and its adjoint on the
// stored species populations; cfinal (optional) receives every C
template <typename T, int SP>
__global__ void k_loss(const void* __restrict__ g, const double* __restrict__ target, long n, int ns, T* lg,
                       double* loss, T* cfinal, const AdjFrames fa, int nx, int ny) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    int4 fr;
    const bool live = i < n && adj_frame(fa, (int)(i % nx), (int)(i / nx), nx, ny, fr);
    for (int s = 0; s < ns && i < n; s++) {
        if (!live) {   // outside every problem of a batch: no loss
#pragma unroll
            for (int q = 0; q < 5; q++) lg[(long)(5 * s + q) * n + i] = (T)0;
            if (cfinal) cfinal[(long)s * n + i] = (T)0;
            continue;
        }
        T C = 0;
#pragma unroll
        for (int q = 0; q < 5; q++) C += PopStore<T, SP>::g(g, (long)(5 * s + q) * n + i);
        const double r = (double)C - target[(long)s * n + i];
        r2 += r * r;
        if (cfinal) cfinal[(long)s * n + i] = C;
#pragma unroll
        for (int q = 0; q < 5; q++) lg[(long)(5 * s + q) * n + i] = (T)(2.0 * r);
    }
    for (int o = 32; o > 0; o >>= 1) r2 += __shfl_down(r2, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, r2);
}

// ---------------------------------------------------------------- reverse --
// interpolation + move: lamX (= lambda of X_{t+1}) is lambda_U; adds the
// derivative of the weights and spreads lambda_U onto lamU (sites).
template <typename T>
__global__ void k_interp_bwd(int nm, const T* __restrict__ X, const T* __restrict__ u, T* lamX, T* lamU, int nx,
                             int ny, const MarkerFrames mf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const MarkerWrap wr(mf, k, nx, ny);
    int sx, sy;
    T hx[4], hy[4], dhx[4], dhy[4];
    stencil<T>(X[2 * k], sx, hx, dhx);
    stencil<T>(X[2 * k + 1], sy, hy, dhy);
    const T lx = lamX[2 * k], ly = lamX[2 * k + 1];
    T gx = 0, gy = 0;
#pragma unroll
    for (int qx = 0; qx < 4; qx++)
#pragma unroll
        for (int qy = 0; qy < 4; qy++) {
            const long s = wr.site(sx + qx, sy + qy);
            const T ul = u[2 * s] * lx + u[2 * s + 1] * ly;
            gx += (dhx[qx] * hy[qy]) / (T)16 * ul;
            gy += (hx[qx] * dhy[qy]) / (T)16 * ul;
            const T d = (hx[qx] * hy[qy]) / (T)16;
            atomicAdd(&lamU[2 * s], d * lx);
            atomicAdd(&lamU[2 * s + 1], d * ly);
        }
    lamX[2 * k] = lx + gx;
    lamX[2 * k + 1] = ly + gy;
}

template <typename T>
struct BwdArgs {
    int nx, ny;
    long n;
    const void* fs;   // f_t, g_t (pre-stream states of the step, in the population storage)
    const void* gs;
    const T* Fs;   // site force of the step
    const T* lf1;  // adjoints of f_{t+1}, g_{t+1}
    const T* lg1;
    T* lamU;       // adjoint of the output velocity from the interpolation (read, cleared)
    T* lf0;        // adjoints of f_t, g_t (written by push)
    T* lg0;
    T* lamF;       // adjoint of the site force
    const uint32_t* dom;
    const double* src;
    T tauI;
    T tauIc[LBIB_MAX_SPECIES], decay[LBIB_MAX_SPECIES];
    T Km[LBIB_MAX_SPECIES * LBIB_MAX_SPECIES];
    int hasK;
    int ncell;
    double* grad;  // [s] tau_c of species s, [ns + c] rate of cell c, [ns + ncell + 1 + s ns + j] K_sj
    AdjFrames fr;
};

// Per-site adjoint of pull + fluid collide + species collides.
template <typename T, int NS, int SP>
__global__ __launch_bounds__(256) void k_lattice_bwd(const BwdArgs<T> a) {
    using PS = PopStore<T, SP>;
    // [ns + ncell + 1 + ns^2] gradient slots, laid out like a.grad
    extern __shared__ double sgrad[];
    const int ng = NS + a.ncell + 1 + NS * NS;
    for (int c = threadIdx.x; c < ng; c += blockDim.x) sgrad[c] = 0.0;
    __syncthreads();
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // this site's gradient terms (zero past the lattice end), summed over the
    // wavefront before one LDS atomic per slot: every lane adds to the same
    // few slots, so per-lane atomics would serialise
    double gt[NS], gK[NS * NS], grate = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) gt[s] = 0.0;
#pragma unroll
    for (int s = 0; s < NS * NS; s++) gK[s] = 0.0;
    uint32_t dmw = 0;
    int4 fr;
    if (i < a.n && adj_frame(a.fr, (int)(i % a.nx), (int)(i / a.nx), a.nx, a.ny, fr)) {
        using Q = Cst<T>;
        const int x = (int)(i % a.nx), y = (int)(i / a.nx);
        long idx[9];
        pull_index(x, y, a.nx, fr, idx);
        const long n = a.n;
        T d[9], g[NS][5];
#pragma unroll
        for (int q = 0; q < 9; q++) d[q] = PS::f(a.fs, q * n + idx[q], q);
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++) g[s][q] = PS::g(a.gs, (5 * s + q) * n + idx[q]);
        const T Fx = a.Fs[2 * i], Fy = a.Fs[2 * i + 1];
        // forward values
        T rho = 0;
#pragma unroll
        for (int q = 0; q < 9; q++) rho += d[q];
        const T rhoI = (T)1 / rho;
        const T ux = (d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW])) * rhoI;
        const T uy = (d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE])) * rhoI;
        const T uox = ux + Fx, uoy = uy + Fy;   // post-collision velocity (exact relation)
        T C[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            C[s] = 0;
#pragma unroll
            for (int q = 0; q < 5; q++) C[s] += g[s][q];
        }
        const uint32_t dm = a.dom[i];
        // adjoints of step t+1 at this site
        T lf[9], lg[NS][5];
#pragma unroll
        for (int q = 0; q < 9; q++) lf[q] = a.lf1[q * n + i];
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++) lg[s][q] = a.lg1[(5 * s + q) * n + i];
        T luox = a.lamU[2 * i], luoy = a.lamU[2 * i + 1];
        a.lamU[2 * i] = 0;
        a.lamU[2 * i + 1] = 0;
        // species s: g'_p = g_p (1 - tauIc_s) + geq_p(C_s, uo) tauIc_s + R_s r_p with
        // R_s = sum_j K_sj C_j + (inside: rate of the cell for s = 0; outside: -decay_s C_s)
        const T ae[5] = {(T)1 / (T)3, ((T)1 + (T)3 * uox) / (T)6, ((T)1 + (T)3 * uoy) / (T)6,
                         ((T)1 - (T)3 * uox) / (T)6, ((T)1 - (T)3 * uoy) / (T)6};
        const T rr[5] = {(T)1 / (T)3, (T)1 / (T)6, (T)1 / (T)6, (T)1 / (T)6, (T)1 / (T)6};
        T P[NS];   // adjoint of R_s
#pragma unroll
        for (int s = 0; s < NS; s++) {
            P[s] = 0;
#pragma unroll
            for (int p = 0; p < 5; p++) P[s] += lg[s][p] * rr[p];
        }
        T lgh[NS][5];
#pragma unroll
        for (int j = 0; j < NS; j++) {
            const T tauIc = a.tauIc[j];
            const T dRdC = dm != 0u ? (T)0 : -a.decay[j];
            T S = 0, gtau = 0;
#pragma unroll
            for (int p = 0; p < 5; p++) {
                S += lg[j][p] * (ae[p] * tauIc + rr[p] * dRdC);
                gtau += lg[j][p] * (g[j][p] - ae[p] * C[j]);
            }
            if (a.hasK) {
#pragma unroll
                for (int s = 0; s < NS; s++) S += P[s] * a.Km[s * LBIB_MAX_SPECIES + j];
            }
#pragma unroll
            for (int q = 0; q < 5; q++) lgh[j][q] = lg[j][q] * ((T)1 - tauIc) + S;
            // d geq_p / d uo = C/2 e_p for p = E, N, W, S
            luox += C[j] / (T)2 * tauIc * (lg[j][DE] - lg[j][DW]);
            luoy += C[j] / (T)2 * tauIc * (lg[j][DN] - lg[j][DS]);
            // d g'_p / d tau_c = (g_p - geq_p) / tau_c^2
            gt[j] = (double)(gtau * tauIc * tauIc);
        }
        // fluid: f'_q = f_q (1 - tauI) + feq_q(rho, u) tauI + 3 w_q rho (e_q . F), uo = u + F
        const int ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
        const int ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
        const T w[9] = {Q::W0, Q::W1, Q::W1, Q::W1, Q::W1, Q::W2, Q::W2, Q::W2, Q::W2};
        const T u2 = ux * ux + uy * uy;
        T Asum = 0, Bx = luox, By = luoy, lFx = luox, lFy = luoy;
#pragma unroll
        for (int p = 0; p < 9; p++) {
            const T eu = (T)ex[p] * ux + (T)ey[p] * uy;
            const T feq_r = w[p] * ((T)1 + (T)3 * eu + (T)4.5 * eu * eu - (T)1.5 * u2);   // feq / rho
            const T eF = (T)ex[p] * Fx + (T)ey[p] * Fy;
            Asum += lf[p] * (a.tauI * feq_r + (T)3 * w[p] * eF);
            const T cf = lf[p] * a.tauI * w[p] * rho;
            Bx += cf * ((T)3 * (T)ex[p] + (T)9 * eu * (T)ex[p] - (T)3 * ux);
            By += cf * ((T)3 * (T)ey[p] + (T)9 * eu * (T)ey[p] - (T)3 * uy);
            lFx += lf[p] * (T)3 * w[p] * rho * (T)ex[p];
            lFy += lf[p] * (T)3 * w[p] * rho * (T)ey[p];
        }
        // push: the adjoint of the pull from x - e_q lands there
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const T lfh = lf[q] * ((T)1 - a.tauI) + Asum + (Bx * ((T)ex[q] - ux) + By * ((T)ey[q] - uy)) * rhoI;
            a.lf0[q * n + idx[q]] = lfh;
        }
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 5; q++) a.lg0[(5 * s + q) * n + idx[q]] = lgh[s][q];
        a.lamF[2 * i] = lFx;
        a.lamF[2 * i + 1] = lFy;
        if (dm != 0u && (int)dm <= a.ncell) {
            dmw = dm;
            grate = (double)P[0];
        }
        if (a.hasK) {
#pragma unroll
            for (int s = 0; s < NS; s++)
#pragma unroll
                for (int j = 0; j < NS; j++) gK[s * NS + j] = (double)(P[s] * C[j]);
        }
    }
    const int lane = (int)threadIdx.x & 63;
    auto wsum = [](double v) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
    };
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double v = wsum(gt[s]);
        if (lane == 0) atomicAdd(&sgrad[s], v);
    }
    if (a.hasK) {
#pragma unroll
        for (int q = 0; q < NS * NS; q++) {
            const double v = wsum(gK[q]);
            if (lane == 0) atomicAdd(&sgrad[NS + a.ncell + 1 + q], v);
        }
    }
    // per-cell rates: one sum when the whole wavefront lies in one cell
    const uint32_t dm0 = (uint32_t)__shfl((int)dmw, 0, 64);
    if (__all(dmw == dm0)) {
        const double v = wsum(grate);
        if (lane == 0 && dm0 != 0u) atomicAdd(&sgrad[NS + dm0], v);
    } else if (dmw != 0u) {
        atomicAdd(&sgrad[NS + dmw], grate);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ng; c += blockDim.x)
        if (sgrad[c] != 0.0) atomicAdd(&a.grad[c], sgrad[c]);
}

// spread adjoint: lambda of the marker force and the weight derivative term
template <typename T>
__global__ void k_spread_bwd(int nm, const T* __restrict__ X, const T* __restrict__ Fm, const T* __restrict__ lamF,
                             T* lamFm, T* lamX, int nx, int ny, const MarkerFrames mf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const MarkerWrap wr(mf, k, nx, ny);
    int sx, sy;
    T hx[4], hy[4], dhx[4], dhy[4];
    stencil<T>(X[2 * k], sx, hx, dhx);
    stencil<T>(X[2 * k + 1], sy, hy, dhy);
    const T fx = Fm[2 * k], fy = Fm[2 * k + 1];
    T lx = 0, ly = 0, gx = 0, gy = 0;
#pragma unroll
    for (int qx = 0; qx < 4; qx++)
#pragma unroll
        for (int qy = 0; qy < 4; qy++) {
            const long s = wr.site(sx + qx, sy + qy);
            const T d = (hx[qx] * hy[qy]) / (T)16;
            const T Lx = lamF[2 * s], Ly = lamF[2 * s + 1];
            lx += d * Lx;
            ly += d * Ly;
            const T fl = fx * Lx + fy * Ly;
            gx += (dhx[qx] * hy[qy]) / (T)16 * fl;
            gy += (hx[qx] * dhy[qy]) / (T)16 * fl;
        }
    lamFm[2 * k] = lx;
    lamFm[2 * k + 1] = ly;
    lamX[2 * k] += gx;
    lamX[2 * k + 1] += gy;
}

// tension adjoint: connection (a, b) gives F_a += kappa n, F_b -= kappa n
// with n = (X_b - X_a)/|X_b - X_a|; each marker gathers its two connections.
// Junction (k, m): F_k += k_j phi(d), d = X_m - X_k, phi(d) = (|d| - l0) d/|d|,
// listed at both ends; J = d phi / d d = (1 - l0/|d|) I + (l0/|d|) n n^T is
// symmetric, so lambda_X_k += k_j J (lambda_F_m - lambda_F_k) and
// dL/dk_j = sum_k lambda_F_k . phi(d).
template <typename T>
__global__ void k_force_bwd(int nm, const T* __restrict__ X, const int32_t* __restrict__ nxt,
                            const int32_t* __restrict__ prv, T kappa, const T* __restrict__ lamFm, T* lamX,
                            double* gkappa, const Junctions jn, double* gkj) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double gk = 0.0, gj = 0.0;
    if (k < nm) {
        const T xk = X[2 * k], yk = X[2 * k + 1];
        T ax = 0, ay = 0;
        // as the first end of (k, next)
        {
            const int b = nxt[k];
            const T dx = X[2 * b] - xk, dy = X[2 * b + 1] - yk;
            const T dist = sqrt(dx * dx + dy * dy);
            const T nxv = dx / dist, nyv = dy / dist;
            const T vx = lamFm[2 * k] - lamFm[2 * b], vy = lamFm[2 * k + 1] - lamFm[2 * b + 1];
            const T nv = nxv * vx + nyv * vy;
            ax -= kappa * (vx - nxv * nv) / dist;
            ay -= kappa * (vy - nyv * nv) / dist;
            gk = (double)nv;
        }
        // as the second end of (prev, k)
        {
            const int p = prv[k];
            const T dx = xk - X[2 * p], dy = yk - X[2 * p + 1];
            const T dist = sqrt(dx * dx + dy * dy);
            const T nxv = dx / dist, nyv = dy / dist;
            const T vx = lamFm[2 * p] - lamFm[2 * k], vy = lamFm[2 * p + 1] - lamFm[2 * k + 1];
            const T nv = nxv * vx + nyv * vy;
            ax += kappa * (vx - nxv * nv) / dist;
            ay += kappa * (vy - nyv * nv) / dist;
        }
        if (jn.ofs) {
            const T kj = (T)jn.kj, l0 = (T)jn.l0;
            const T lkx = lamFm[2 * k], lky = lamFm[2 * k + 1];
            for (int p = jn.ofs[k]; p < jn.ofs[k + 1]; p++) {
                const int m = jn.part[p];
                const T dx = X[2 * m] - xk, dy = X[2 * m + 1] - yk;
                const T dist = sqrt(dx * dx + dy * dy);
                const T nxv = dx / dist, nyv = dy / dist;
                const T vx = lamFm[2 * m] - lkx, vy = lamFm[2 * m + 1] - lky;
                const T a1 = (T)1 - l0 / dist, a2 = l0 / dist * (nxv * vx + nyv * vy);
                ax += kj * (a1 * vx + a2 * nxv);
                ay += kj * (a1 * vy + a2 * nyv);
                gj += (double)((dist - l0) * (lkx * nxv + lky * nyv));
            }
        }
        lamX[2 * k] += ax;
        lamX[2 * k + 1] += ay;
    }
    for (int o = 32; o > 0; o >>= 1) {
        gk += __shfl_down(gk, o, 64);
        gj += __shfl_down(gj, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(gkappa, gk);
        if (jn.ofs) atomicAdd(gkj, gj);
    }
}

}  // namespace adj
}  // namespace lbib

using namespace lbib;

struct lbib_adj {
    lbib_adj_config cfg;
    long n = 0;
    int nm = 0, ncell = 0, K = 50, prec = LBIB_FP32;
    size_t tb = 4;                  // bytes of the arithmetic type
    int sprec = LBIB_FP32;          // population storage (the arithmetic type, or 16-bit deviations)
    size_t sb = 4;                  // bytes of one stored population value
    int scheme = 0;                 // 0: checkpoint every K steps + segment store; 1: binomial (revolve)
    int snaps = 0;                  // revolve: snapshot slots
    hipStream_t stream = nullptr;
    std::string err;
    // state: populations in the storage type, markers etc. in the arithmetic type
    void *f[2] = {nullptr, nullptr}, *g[2] = {nullptr, nullptr};
    void *X = nullptr, *Xs = nullptr, *Fm = nullptr, *Fs = nullptr, *u = nullptr;
    // problem
    void *f0 = nullptr, *g0 = nullptr, *X0 = nullptr;
    int32_t *nxt = nullptr, *prv = nullptr;
    uint32_t* dom = nullptr;
    double *target = nullptr, *src = nullptr, *grad = nullptr, *loss = nullptr;
    // checkpoints (or revolve snapshots) and the segment store
    void* ck = nullptr;             // nck state slots
    int nck = 0;
    void* seg = nullptr;            // K states + K x (4 n + 2 nm) forces and velocities
    // adjoints
    void *lf[2] = {nullptr, nullptr}, *lg[2] = {nullptr, nullptr};
    void *lX = nullptr, *lFm = nullptr, *lU = nullptr, *lF = nullptr;
    int steps = 0;
    double kappa = 0.01;
    std::vector<double> rates;
    // species (lbib_adj_set_species; one by default) and the reaction matrix
    int ns = 1;
    double tau_c[LBIB_MAX_SPECIES] = {0.8, 0.8, 0.8, 0.8};
    double decay[LBIB_MAX_SPECIES] = {0, 0, 0, 0};
    double Km[LBIB_MAX_SPECIES * LBIB_MAX_SPECIES] = {};
    void* cfin = nullptr;           // final concentrations [ns][n] (on request)
    // junction springs (lbib_adj_set_junctions): partners of each marker
    int32_t *jofs = nullptr, *jpart = nullptr;
    double kj = 0.0, l0 = 0.5;
    // batch of problems (lbib_adj_set_batch_layout)
    int32_t* rtile = nullptr;
    int4* rects = nullptr;
    int32_t* mrect = nullptr;   // rectangle of every marker (batches)
    int ntx = 0;
    std::vector<double> xy_host;    // initial marker positions (layout checks)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double fwd_ms = 0.0, grad_ms = 0.0;
    long fwd_steps_run = 0, grad_steps_run = 0;   // forward steps taken by the last forward / gradient
    bool have_problem = false, have_forward = false;
};

namespace {

#define ADJ_HIP(call)                                                      \
    do {                                                                   \
        hipError_t e_ = (call);                                            \
        if (e_ != hipSuccess) {                                            \
            a->err = std::string(#call) + ": " + hipGetErrorString(e_);   \
            return -1;                                                     \
        }                                                                  \
    } while (0)

int adj_alloc(lbib_adj* a, void** p, size_t bytes) {
    *p = nullptr;
    ADJ_HIP(hipMalloc(p, std::max<size_t>(bytes, 16)));
    ADJ_HIP(hipMemsetAsync(*p, 0, std::max<size_t>(bytes, 16), a->stream));
    return 0;
}
void adj_free(void*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
inline int nb(long count) { return (int)((count + 255) / 256); }
adj::MarkerFrames mframes(const lbib_adj* a) { return adj::MarkerFrames{a->rtile ? a->mrect : nullptr, a->rects}; }

template <typename T>
T* tp(void* p) { return static_cast<T*>(p); }

// state slots (bytes): fluid populations, species populations (storage
// type), then the markers (arithmetic type) at a 256-byte boundary
inline size_t gplanes(const lbib_adj* a) { return 5 * (size_t)a->ns; }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
inline size_t fbytes(const lbib_adj* a) { return 9 * (size_t)a->n * a->sb; }
inline size_t gbytes(const lbib_adj* a) { return gplanes(a) * (size_t)a->n * a->sb; }
inline size_t xbytes(const lbib_adj* a) { return 2 * (size_t)a->nm * a->tb; }
inline size_t xoff(const lbib_adj* a) { return align256(fbytes(a) + gbytes(a)); }
inline size_t state_bytes(const lbib_adj* a) { return align256(xoff(a) + xbytes(a)); }
// gradient slots: tau_c [ns], rates [ncell + 1], K [ns^2], then kappa and k_j
inline int grad_len(int ns, int ncell) { return ns + ncell + 1 + ns * ns + 2; }
inline size_t extra_bytes(const lbib_adj* a) { return align256((4 * (size_t)a->n + 2 * (size_t)a->nm) * a->tb); }

template <typename T>
struct Slot {
    void *f, *g;   // 9n, 5 ns n stored values
    T* X;          // 2nm
};
template <typename T>
Slot<T> state_slot(void* base, const lbib_adj* a, int j) {
    char* s = static_cast<char*>(base) + (size_t)j * state_bytes(a);
    return {s, s + fbytes(a), reinterpret_cast<T*>(s + xoff(a))};
}
template <typename T>
struct Extra {
    T *Fs, *u, *Fm;   // 2n, 2n, 2nm
};
template <typename T>
Extra<T> extra_slot(lbib_adj* a, int j) {
    T* s = reinterpret_cast<T*>(static_cast<char*>(a->seg) + (size_t)a->K * state_bytes(a) +
                                (size_t)j * extra_bytes(a));
    return {s, s + 2 * a->n, s + 4 * a->n};
}
template <typename T>
int copy_state(lbib_adj* a, const Slot<T>& dst, const Slot<T>& srcs) {
    ADJ_HIP(hipMemcpyAsync(dst.f, srcs.f, fbytes(a), hipMemcpyDeviceToDevice, a->stream));
    ADJ_HIP(hipMemcpyAsync(dst.g, srcs.g, gbytes(a), hipMemcpyDeviceToDevice, a->stream));
    if (a->nm > 0) ADJ_HIP(hipMemcpyAsync(dst.X, srcs.X, xbytes(a), hipMemcpyDeviceToDevice, a->stream));
    return 0;
}

template <typename T>
adj::LatArgs<T> lat_args(lbib_adj* a, const void* fs, const void* gs, void* fd, void* gd, const T* Fs, T* u) {
    adj::LatArgs<T> l;
    l.nx = a->cfg.nx; l.ny = a->cfg.ny; l.n = a->n;
    l.fs = fs; l.gs = gs; l.fd = fd; l.gd = gd; l.Fs = Fs; l.u = u;
    l.dom = a->dom; l.src = a->src;
    l.tauI = (T)(1.0 / a->cfg.tau_fluid);
    l.fr = adj::AdjFrames{a->rtile, a->rects, a->ntx};
    l.hasK = 0;
    for (int s = 0; s < LBIB_MAX_SPECIES; s++) {
        l.tauIc[s] = (T)(1.0 / a->tau_c[s]);
        l.decay[s] = (T)a->decay[s];
        for (int j = 0; j < LBIB_MAX_SPECIES; j++) {
            l.Km[s * LBIB_MAX_SPECIES + j] = (T)a->Km[s * LBIB_MAX_SPECIES + j];
            if (a->Km[s * LBIB_MAX_SPECIES + j] != 0.0) l.hasK = 1;
        }
    }
    return l;
}

template <typename T>
adj::Junctions junctions(const lbib_adj* a) {
    return adj::Junctions{a->jofs, a->jpart, a->kj, a->l0};
}

// the storage variants: the arithmetic type itself, or (FP32 arithmetic
// only) 16-bit deviations
template <typename T, template <typename, int> class Launch, typename... Args>
void by_storage(lbib_adj* a, Args&&... args) {
    if constexpr (std::is_same<T, float>::value) {
        if (a->sprec == LBIB_FP16_DEV) { Launch<T, LBIB_FP16_DEV>::go(a, args...); return; }
        if (a->sprec == LBIB_BF16_DEV) { Launch<T, LBIB_BF16_DEV>::go(a, args...); return; }
    }
    Launch<T, 0>::go(a, args...);
}
template <typename T, int SP>
struct LatLaunch {
    static void go(lbib_adj* a, const adj::LatArgs<T>& l) {
        switch (a->ns) {
            case 1: adj::k_lattice<T, 1, SP><<<nb(a->n), 256, 0, a->stream>>>(l); break;
            case 2: adj::k_lattice<T, 2, SP><<<nb(a->n), 256, 0, a->stream>>>(l); break;
            case 3: adj::k_lattice<T, 3, SP><<<nb(a->n), 256, 0, a->stream>>>(l); break;
            default: adj::k_lattice<T, 4, SP><<<nb(a->n), 256, 0, a->stream>>>(l); break;
        }
    }
};
template <typename T, int SP>
struct BwdLaunch {
    static void go(lbib_adj* a, const adj::BwdArgs<T>& b) {
        const size_t lds = (size_t)(a->ns + a->ncell + 1 + a->ns * a->ns) * sizeof(double);
        switch (a->ns) {
            case 1: adj::k_lattice_bwd<T, 1, SP><<<nb(a->n), 256, lds, a->stream>>>(b); break;
            case 2: adj::k_lattice_bwd<T, 2, SP><<<nb(a->n), 256, lds, a->stream>>>(b); break;
            case 3: adj::k_lattice_bwd<T, 3, SP><<<nb(a->n), 256, lds, a->stream>>>(b); break;
            default: adj::k_lattice_bwd<T, 4, SP><<<nb(a->n), 256, lds, a->stream>>>(b); break;
        }
    }
};
template <typename T, int SP>
struct LossLaunch {
    static void go(lbib_adj* a, const void* g, T* cf) {
        adj::k_loss<T, SP><<<nb(a->n), 256, 0, a->stream>>>(g, a->target, a->n, a->ns, tp<T>(a->lg[0]), a->loss, cf,
                                                            adj::AdjFrames{a->rtile, a->rects, a->ntx}, a->cfg.nx,
                                                            a->cfg.ny);
    }
};

// one forward step from (fs, gs, X) into (fd, gd, X); Fs, Fm, u receive the
// step's site force, marker forces and velocity
template <typename T>
int step_fwd(lbib_adj* a, const void* fs, const void* gs, void* fd, void* gd, T* X, T* Fs, T* Fm, T* u) {
    const int nx = a->cfg.nx, ny = a->cfg.ny;
    ADJ_HIP(hipMemsetAsync(Fs, 0, 2 * a->n * sizeof(T), a->stream));
    if (a->nm > 0) {
        adj::k_force<T><<<nb(a->nm), 256, 0, a->stream>>>(a->nm, X, a->nxt, a->prv, (T)a->kappa, Fm,
                                                          junctions<T>(a));
        adj::k_spread<T><<<nb(a->nm), 256, 0, a->stream>>>(a->nm, X, Fm, Fs, nx, ny, mframes(a));
    }
    by_storage<T, LatLaunch>(a, lat_args<T>(a, fs, gs, fd, gd, Fs, u));
    if (a->nm > 0) adj::k_interp_move<T><<<nb(a->nm), 256, 0, a->stream>>>(a->nm, X, u, nx, ny, mframes(a));
    ADJ_HIP(hipGetLastError());
    return 0;
}

// ---- binomial checkpointing (revolve) ----------------------------------------
// beta(s, r) = C(s + r, s): the most steps s snapshots reverse when every
// step is recomputed at most r times (Griewank & Walther, Algorithm 799),
// saturated far above any step count.
double rev_beta(int s, int r) {
    double b = 1.0;
    for (int k = 1; k <= s; k++) b = b * (double)(r + k) / (double)k;
    return b;
}
// First snapshot position (steps after t0) for l steps with s free slots:
// the least r with beta(s, r) >= l, then the smallest advance that leaves
// the right part within beta(s - 1, r) (the left part then fits beta(s, r - 1)).
int rev_split(int l, int s) {
    int r = 0;
    while (rev_beta(s, r) < (double)l) r++;
    double m = (double)l - rev_beta(s - 1, r);
    if (m < 1.0) m = 1.0;
    const double cap = rev_beta(s, r - 1);
    if (m > cap) m = cap;
    if (m > l - 1) m = l - 1;
    return (int)m;
}
// snapshot positions the forward sweep stores: the right spine of the
// recursion (slot k holds the state before step pos[k])
std::vector<int> rev_spine(int steps, int s) {
    std::vector<int> pos;
    int t0 = 0, free = s;
    while (free > 0 && steps - t0 > 1) {
        t0 += rev_split(steps - t0, free);
        pos.push_back(t0);
        free--;
    }
    return pos;
}

template <typename T>
int forward_t(lbib_adj* a, int steps, double* loss, double* cfinal) {
    const std::vector<int> spine = a->scheme == 1 ? rev_spine(steps, a->snaps) : std::vector<int>();
    const int need = a->scheme == 1 ? std::max(a->snaps, 1) : steps / a->K + 1;
    if (need > a->nck) {
        adj_free(a->ck);
        a->nck = 0;
        if (adj_alloc(a, &a->ck, (size_t)need * state_bytes(a))) return -1;
        a->nck = need;
    }
    ADJ_HIP(hipMemcpyAsync(a->f[0], a->f0, fbytes(a), hipMemcpyDeviceToDevice, a->stream));
    ADJ_HIP(hipMemcpyAsync(a->g[0], a->g0, gbytes(a), hipMemcpyDeviceToDevice, a->stream));
    if (a->nm > 0) ADJ_HIP(hipMemcpyAsync(a->X, a->X0, xbytes(a), hipMemcpyDeviceToDevice, a->stream));
    ADJ_HIP(hipEventRecord(a->e0, a->stream));
    int cur = 0;
    size_t sp = 0;
    for (int t = 0; t < steps; t++) {
        int slot = -1;
        if (a->scheme == 0 && t % a->K == 0) slot = t / a->K;
        if (a->scheme == 1 && sp < spine.size() && spine[sp] == t) slot = (int)sp++;
        if (slot >= 0) {
            const Slot<T> w{a->f[cur], a->g[cur], tp<T>(a->X)};
            if (copy_state<T>(a, state_slot<T>(a->ck, a, slot), w)) return -1;
        }
        if (step_fwd<T>(a, a->f[cur], a->g[cur], a->f[1 - cur], a->g[1 - cur], tp<T>(a->X), tp<T>(a->Fs),
                        tp<T>(a->Fm), tp<T>(a->u)))
            return -1;
        cur = 1 - cur;
    }
    ADJ_HIP(hipMemsetAsync(a->loss, 0, sizeof(double), a->stream));
    T* cf = nullptr;
    if (cfinal) {
        if (!a->cfin && adj_alloc(a, &a->cfin, (size_t)a->ns * a->n * sizeof(T))) return -1;
        cf = tp<T>(a->cfin);
    }
    by_storage<T, LossLaunch>(a, (const void*)a->g[cur], cf);
    ADJ_HIP(hipEventRecord(a->e1, a->stream));
    ADJ_HIP(hipEventSynchronize(a->e1));
    float ms = 0.0f;
    ADJ_HIP(hipEventElapsedTime(&ms, a->e0, a->e1));
    a->fwd_ms = ms;
    a->fwd_steps_run = steps;
    ADJ_HIP(hipMemcpy(loss, a->loss, sizeof(double), hipMemcpyDeviceToHost));
    if (cfinal) {
        const size_t cnt = (size_t)a->ns * a->n;
        std::vector<T> tmp(cnt);
        ADJ_HIP(hipMemcpy(tmp.data(), cf, cnt * sizeof(T), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cnt; i++) cfinal[i] = (double)tmp[i];
    }
    a->steps = steps;
    return 0;
}

// The reverse of one step: state t in st (pre-move markers), its forces and
// velocity in e; lambda of t+1 in lf/lg[cur] becomes lambda of t in lf/lg[1-cur].
template <typename T>
int reverse_step(lbib_adj* a, const Slot<T>& st, const Extra<T>& e, int& cur) {
    const int nx = a->cfg.nx, ny = a->cfg.ny;
    const int gl = grad_len(a->ns, a->ncell);
    if (a->nm > 0)
        adj::k_interp_bwd<T><<<nb(a->nm), 256, 0, a->stream>>>(a->nm, st.X, e.u, tp<T>(a->lX), tp<T>(a->lU), nx, ny,
                                                                 mframes(a));
    adj::BwdArgs<T> b;
    b.nx = nx; b.ny = ny; b.n = a->n;
    b.fs = st.f; b.gs = st.g; b.Fs = e.Fs;
    b.lf1 = tp<T>(a->lf[cur]); b.lg1 = tp<T>(a->lg[cur]);
    b.lamU = tp<T>(a->lU);
    b.lf0 = tp<T>(a->lf[1 - cur]); b.lg0 = tp<T>(a->lg[1 - cur]);
    b.lamF = tp<T>(a->lF);
    b.dom = a->dom; b.src = a->src;
    const adj::LatArgs<T> l = lat_args<T>(a, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    b.tauI = l.tauI;
    for (int q = 0; q < LBIB_MAX_SPECIES; q++) { b.tauIc[q] = l.tauIc[q]; b.decay[q] = l.decay[q]; }
    for (int q = 0; q < LBIB_MAX_SPECIES * LBIB_MAX_SPECIES; q++) b.Km[q] = l.Km[q];
    b.hasK = l.hasK;
    b.fr = l.fr;
    b.ncell = a->ncell;
    b.grad = a->grad;
    by_storage<T, BwdLaunch>(a, b);
    if (a->nm > 0) {
        adj::k_spread_bwd<T><<<nb(a->nm), 256, 0, a->stream>>>(a->nm, st.X, e.Fm, tp<T>(a->lF), tp<T>(a->lFm),
                                                                 tp<T>(a->lX), nx, ny, mframes(a));
        adj::k_force_bwd<T><<<nb(a->nm), 256, 0, a->stream>>>(a->nm, st.X, a->nxt, a->prv, (T)a->kappa,
                                                                tp<T>(a->lFm), tp<T>(a->lX), a->grad + gl - 2,
                                                                junctions<T>(a), a->grad + gl - 1);
    }
    ADJ_HIP(hipGetLastError());
    cur = 1 - cur;
    return 0;
}

// Revolve.  The current state is a view: a snapshot slot (or the initial
// state, slot -1), read in place, or one of the working buffers f/g[0..1]
// after steps were taken from it.  Markers move in place, so a view of a
// slot first copies its markers to X (2 nm values, not the populations).
template <typename T>
struct Revolver {
    lbib_adj* a;
    int cur = 0;
    Slot<T> st{};   // current state
    int w = -1;     // working buffer holding st, -1: a slot
    Slot<T> work(int k) const { return {a->f[k], a->g[k], tp<T>(a->X)}; }
    Slot<T> source(int slot) const {
        return slot < 0 ? Slot<T>{a->f0, a->g0, tp<T>(a->X0)} : state_slot<T>(a->ck, a, slot);
    }
    void view(int slot) {
        st = source(slot);
        w = -1;
    }
    // working buffer that does not hold the current state
    int other() const { return w == 0 ? 1 : 0; }
    int advance(int nsteps) {
        for (int k = 0; k < nsteps; k++) {
            if (w < 0 && a->nm > 0)
                ADJ_HIP(hipMemcpyAsync(a->X, st.X, xbytes(a), hipMemcpyDeviceToDevice, a->stream));
            const int d = other();
            if (step_fwd<T>(a, st.f, st.g, a->f[d], a->g[d], tp<T>(a->X), tp<T>(a->Fs), tp<T>(a->Fm), tp<T>(a->u)))
                return -1;
            st = work(d);
            w = d;
            a->grad_steps_run++;
        }
        return 0;
    }
    // reverse of the step from the current state: recompute it into the
    // other working buffer with the markers moved in Xs, st.X staying at t
    int reverse_one() {
        if (a->nm > 0) ADJ_HIP(hipMemcpyAsync(a->Xs, st.X, xbytes(a), hipMemcpyDeviceToDevice, a->stream));
        const int d = other();
        if (step_fwd<T>(a, st.f, st.g, a->f[d], a->g[d], tp<T>(a->Xs), tp<T>(a->Fs), tp<T>(a->Fm), tp<T>(a->u)))
            return -1;
        a->grad_steps_run++;
        const Extra<T> e{tp<T>(a->Fs), tp<T>(a->u), tp<T>(a->Fm)};
        return reverse_step<T>(a, st, e, cur);
    }
    // adjoint from t1 back to t0, the state of t0 being in `slot`; `free`
    // snapshot slots from a->snaps - free up; on the right spine the
    // forward sweep already stored the snapshots
    int run(int t0, int slot, int t1, int free, bool spine) {
        const int l = t1 - t0;
        if (l == 1) {
            view(slot);
            return reverse_one();
        }
        if (free == 0) {
            for (int t = t1 - 1; t >= t0; t--) {
                view(slot);
                if (advance(t - t0) || reverse_one()) return -1;
            }
            return 0;
        }
        const int m = t0 + rev_split(l, free);
        const int k = a->snaps - free;
        if (!spine) {
            view(slot);
            if (advance(m - t0) || copy_state<T>(a, state_slot<T>(a->ck, a, k), st)) return -1;
        }
        if (run(m, k, t1, free - 1, spine)) return -1;
        return run(t0, slot, m, free, false);
    }
};

// gradient slots into the caller's arrays (any may be null): tau_c [ns],
// kappa, k_j, K [ns ns] row-major, rates [ncell + 1] of species 0 (entry 0 = 0)
template <typename T>
int gradient_t(lbib_adj* a, double* d_tau, double* d_kappa, double* d_kj, double* d_K, double* d_rates) {
    const int steps = a->steps;
    ADJ_HIP(hipEventRecord(a->e0, a->stream));
    a->grad_steps_run = 0;
    // lambda at T: lf = 0, lg from the loss (k_loss wrote lg[0]), lX = 0
    int cur = 0;
    ADJ_HIP(hipMemsetAsync(a->lf[0], 0, 9 * a->n * sizeof(T), a->stream));
    ADJ_HIP(hipMemsetAsync(a->lX, 0, 2 * (size_t)a->nm * sizeof(T), a->stream));
    ADJ_HIP(hipMemsetAsync(a->lU, 0, 2 * a->n * sizeof(T), a->stream));
    const int gl = grad_len(a->ns, a->ncell);
    ADJ_HIP(hipMemsetAsync(a->grad, 0, gl * sizeof(double), a->stream));
    if (a->scheme == 1) {
        Revolver<T> rv{a};
        rv.cur = 0;
        if (rv.run(0, -1, steps, a->snaps, true)) return -1;
    } else {
        const int nseg = (steps + a->K - 1) / a->K;
        for (int sgi = nseg - 1; sgi >= 0; sgi--) {
            const int t0 = sgi * a->K, len = std::min(a->K, steps - t0);
            // recompute the segment: slot j holds the state before step t0 + j
            if (copy_state<T>(a, state_slot<T>(a->seg, a, 0), state_slot<T>(a->ck, a, sgi))) return -1;
            for (int j = 0; j < len; j++) {
                Slot<T> st = state_slot<T>(a->seg, a, j);
                Extra<T> e = extra_slot<T>(a, j);
                // next state: the following slot, or scratch after the segment's last step
                const bool last = j + 1 == len;
                const Slot<T> nx = last ? Slot<T>{a->f[0], a->g[0], tp<T>(a->X)} : state_slot<T>(a->seg, a, j + 1);
                if (a->nm > 0)
                    ADJ_HIP(hipMemcpyAsync(nx.X, st.X, xbytes(a), hipMemcpyDeviceToDevice, a->stream));
                if (step_fwd<T>(a, st.f, st.g, nx.f, nx.g, nx.X, e.Fs, e.Fm, e.u)) return -1;
                a->grad_steps_run++;
            }
            // reverse over the segment
            for (int j = len - 1; j >= 0; j--)
                if (reverse_step<T>(a, state_slot<T>(a->seg, a, j), extra_slot<T>(a, j), cur)) return -1;
        }
    }
    ADJ_HIP(hipEventRecord(a->e1, a->stream));
    ADJ_HIP(hipEventSynchronize(a->e1));
    float ms = 0.0f;
    ADJ_HIP(hipEventElapsedTime(&ms, a->e0, a->e1));
    a->grad_ms = ms;
    std::vector<double> gr(gl);
    ADJ_HIP(hipMemcpy(gr.data(), a->grad, gr.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int ns = a->ns;
    if (d_tau)
        for (int s = 0; s < ns; s++) d_tau[s] = gr[s];
    if (d_rates) {
        d_rates[0] = 0.0;
        for (int c = 1; c <= a->ncell; c++) d_rates[c] = gr[ns + c];
    }
    if (d_K)
        for (int q = 0; q < ns * ns; q++) d_K[q] = gr[ns + a->ncell + 1 + q];
    if (d_kappa) *d_kappa = gr[gl - 2];
    if (d_kj) *d_kj = gr[gl - 1];
    return 0;
}

template <typename T>
int upload_T(lbib_adj* a, void* dst, const double* src, size_t count) {
    std::vector<T> tmp(count);
    for (size_t k = 0; k < count; k++) tmp[k] = (T)src[k];
    ADJ_HIP(hipMemcpy(dst, tmp.data(), count * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// 16-bit deviation of one value (host side of Codec<SP>::enc)
uint16_t encode16(int sp, float v) {
    if (sp == LBIB_FP16_DEV) {
        const _Float16 h = (_Float16)v;
        uint16_t r;
        std::memcpy(&r, &h, 2);
        return r;
    }
    uint32_t u;
    std::memcpy(&u, &v, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
// populations into the storage: planes of n values; fluid planes store f - w_q
int upload_pop(lbib_adj* a, void* dst, const double* src, size_t planes, bool fluid) {
    const size_t count = planes * (size_t)a->n;
    if (a->sprec == a->prec)
        return a->prec == LBIB_FP64 ? upload_T<double>(a, dst, src, count) : upload_T<float>(a, dst, src, count);
    std::vector<uint16_t> tmp(count);
    for (size_t k = 0; k < count; k++) {
        const int q = (int)(k / (size_t)a->n);
        const float w = !fluid ? 0.0f : (q == 0 ? 4.0f / 9.0f : (q < 5 ? 1.0f / 9.0f : 1.0f / 36.0f));
        tmp[k] = encode16(a->sprec, (float)src[k] - w);
    }
    ADJ_HIP(hipMemcpy(dst, tmp.data(), count * 2, hipMemcpyHostToDevice));
    return 0;
}
// population buffers of the storage type (f, g ping-pong, initial state) and
// the store of the checkpointing scheme
int alloc_states(lbib_adj* a) {
    const size_t n = (size_t)a->n;
    for (int b = 0; b < 2; b++) {
        adj_free(a->f[b]);
        adj_free(a->g[b]);
        if (adj_alloc(a, &a->f[b], 9 * n * a->sb) || adj_alloc(a, &a->g[b], gplanes(a) * n * a->sb)) return -1;
    }
    adj_free(a->f0);
    adj_free(a->g0);
    adj_free(a->seg);
    adj_free(a->ck);
    a->nck = 0;
    if (adj_alloc(a, &a->f0, 9 * n * a->sb) || adj_alloc(a, &a->g0, gplanes(a) * n * a->sb)) return -1;
    if (a->scheme == 0 && adj_alloc(a, &a->seg, (size_t)a->K * (state_bytes(a) + extra_bytes(a)))) return -1;
    return 0;
}

}  // namespace

extern "C" {

int lbib_adj_create(const lbib_adj_config* cfg, lbib_adj** out) {
    if (!cfg || !out) return -1;
    *out = nullptr;
    lbib_adj* a = new (std::nothrow) lbib_adj();
    if (!a) return -1;
    *out = a;
    a->cfg = *cfg;
    if (cfg->nx < 8 || cfg->ny < 8 || cfg->n_markers < 0 || cfg->n_cells < 0 || cfg->n_cells > 4096 ||
        !(cfg->tau_fluid > 0.5) || (cfg->precision != LBIB_FP32 && cfg->precision != LBIB_FP64) ||
        cfg->checkpoint_every < 1) {
        a->err = "invalid adjoint configuration (nx, ny >= 8, FP32 or FP64, tau_fluid > 0.5, checkpoint_every >= 1)";
        return -1;
    }
    a->n = (long)cfg->nx * cfg->ny;
    a->nm = cfg->n_markers;
    a->ncell = cfg->n_cells;
    a->K = cfg->checkpoint_every;
    a->prec = cfg->precision;
    a->tb = a->prec == LBIB_FP64 ? 8 : 4;
    a->rates.assign(a->ncell + 1, 0.0);
    a->decay[0] = cfg->decay;
    ADJ_HIP(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
    ADJ_HIP(hipEventCreate(&a->e0));
    ADJ_HIP(hipEventCreate(&a->e1));
    a->sprec = a->prec;
    a->sb = a->tb;
    const size_t tb = a->tb, n = (size_t)a->n, nm = (size_t)a->nm;
    for (int b = 0; b < 2; b++) {
        if (adj_alloc(a, &a->lf[b], 9 * n * tb) || adj_alloc(a, &a->lg[b], 5 * n * tb)) return -1;
    }
    if (alloc_states(a)) return -1;
    if (adj_alloc(a, &a->X, 2 * nm * tb) || adj_alloc(a, &a->Xs, 2 * nm * tb) || adj_alloc(a, &a->Fm, 2 * nm * tb) ||
        adj_alloc(a, &a->Fs, 2 * n * tb) || adj_alloc(a, &a->u, 2 * n * tb) ||
        adj_alloc(a, &a->X0, 2 * nm * tb) || adj_alloc(a, (void**)&a->nxt, nm * 4) ||
        adj_alloc(a, (void**)&a->prv, nm * 4) || adj_alloc(a, (void**)&a->dom, n * 4) ||
        adj_alloc(a, (void**)&a->target, n * 8) || adj_alloc(a, (void**)&a->src, (a->ncell + 1) * 8) ||
        adj_alloc(a, (void**)&a->grad, grad_len(LBIB_MAX_SPECIES, a->ncell) * 8) || adj_alloc(a, (void**)&a->loss, 8) ||
        adj_alloc(a, &a->lX, 2 * nm * tb) || adj_alloc(a, &a->lFm, 2 * nm * tb) || adj_alloc(a, &a->lU, 2 * n * tb) ||
        adj_alloc(a, &a->lF, 2 * n * tb))
        return -1;
    ADJ_HIP(hipStreamSynchronize(a->stream));
    return 0;
}

void lbib_adj_destroy(lbib_adj* a) {
    if (!a) return;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    void** bufs[] = {&a->f[0], &a->f[1], &a->g[0], &a->g[1], &a->lf[0], &a->lf[1], &a->lg[0], &a->lg[1],
                     &a->X, &a->Xs, &a->Fm, &a->Fs, &a->u, &a->f0, &a->g0, &a->X0, &a->ck, &a->seg, &a->lX, &a->lFm,
                     &a->lU, &a->lF, &a->cfin};
    for (void** b : bufs) adj_free(*b);
    void* p[] = {a->nxt, a->prv, a->dom, a->target, a->src, a->grad, a->loss, a->jofs, a->jpart, a->rtile, a->rects, a->mrect};
    for (void* q : p)
        if (q) (void)hipFree(q);
    if (a->e0) (void)hipEventDestroy(a->e0);
    if (a->e1) (void)hipEventDestroy(a->e1);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
}

const char* lbib_adj_last_error(const lbib_adj* a) { return a ? a->err.c_str() : "null adjoint context"; }

int lbib_adj_set_problem(lbib_adj* a, const double* f0, const double* g0, const double* xy0, const int32_t* next,
                         const uint32_t* dom, const double* target) {
    if (!a || !f0 || !g0 || !dom || !target || (a->nm > 0 && (!xy0 || !next))) return -1;
    if (a->ns != 1) {
        a->err = "lbib_adj_set_problem: set the problem before lbib_adj_set_species";
        return -1;
    }
    const size_t n = (size_t)a->n, nm = (size_t)a->nm;
    // ring successors must form closed rings (every marker has one predecessor)
    std::vector<int32_t> prv(nm, -1);
    for (size_t k = 0; k < nm; k++) {
        if (next[k] < 0 || next[k] >= (int32_t)nm || next[k] == (int32_t)k || prv[next[k]] >= 0) {
            a->err = "lbib_adj_set_problem: next[] must describe closed rings of at least two markers";
            return -1;
        }
        prv[next[k]] = (int32_t)k;
    }
    for (size_t i = 0; i < n; i++)
        if ((int)dom[i] > a->ncell) {
            a->err = "lbib_adj_set_problem: cell map entry above n_cells";
            return -1;
        }
    // markers start on the periodic lattice (their stencils wrap; a ring must
    // not straddle the periodic edge, since the tension uses plain differences)
    for (size_t k = 0; k < nm; k++)
        if (!(xy0[2 * k] >= 0.0 && xy0[2 * k] < a->cfg.nx && xy0[2 * k + 1] >= 0.0 && xy0[2 * k + 1] < a->cfg.ny)) {
            a->err = "lbib_adj_set_problem: markers must start inside the lattice";
            return -1;
        }
    // a new problem drops the batch layout (its per-marker rectangles refer to
    // the old positions): set it again afterwards
    ADJ_HIP(hipStreamSynchronize(a->stream));
    for (int32_t** q : {&a->rtile, &a->mrect}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    if (a->rects) (void)hipFree(a->rects);
    a->rects = nullptr;
    const bool f64 = a->prec == LBIB_FP64;
    if (upload_pop(a, a->f0, f0, 9, true) || upload_pop(a, a->g0, g0, 5, false)) return -1;
    if (nm > 0) {
        if (f64 ? upload_T<double>(a, a->X0, xy0, 2 * nm) : upload_T<float>(a, a->X0, xy0, 2 * nm)) return -1;
        ADJ_HIP(hipMemcpy(a->nxt, next, nm * 4, hipMemcpyHostToDevice));
        ADJ_HIP(hipMemcpy(a->prv, prv.data(), nm * 4, hipMemcpyHostToDevice));
    }
    ADJ_HIP(hipMemcpy(a->dom, dom, n * 4, hipMemcpyHostToDevice));
    ADJ_HIP(hipMemcpy(a->target, target, n * 8, hipMemcpyHostToDevice));
    a->xy_host.assign(xy0 ? xy0 : (const double*)nullptr, xy0 ? xy0 + 2 * nm : (const double*)nullptr);
    a->have_problem = true;
    a->have_forward = false;
    return 0;
}

int lbib_adj_set_species(lbib_adj* a, int nspecies, const double* g0, const double* target, const double* decay) {
    if (!a || !g0 || !target || !decay) return -1;
    if (!a->have_problem) { a->err = "lbib_adj_set_species: set the problem first"; return -1; }
    if (nspecies < 1 || nspecies > LBIB_MAX_SPECIES) {
        a->err = "lbib_adj_set_species: 1 to LBIB_MAX_SPECIES species";
        return -1;
    }
    ADJ_HIP(hipStreamSynchronize(a->stream));
    a->ns = nspecies;
    const size_t tb = a->tb, n = (size_t)a->n, gp = gplanes(a);
    for (int b = 0; b < 2; b++) {
        adj_free(a->g[b]);
        adj_free(a->lg[b]);
        if (adj_alloc(a, &a->g[b], gp * n * a->sb) || adj_alloc(a, &a->lg[b], gp * n * tb)) return -1;
    }
    adj_free(a->g0);
    adj_free(a->seg);
    adj_free(a->ck);
    adj_free(a->cfin);
    a->nck = 0;
    if (a->target) (void)hipFree(a->target);
    a->target = nullptr;
    if (adj_alloc(a, &a->g0, gp * n * a->sb) || adj_alloc(a, (void**)&a->target, nspecies * n * 8) ||
        (a->scheme == 0 && adj_alloc(a, &a->seg, (size_t)a->K * (state_bytes(a) + extra_bytes(a)))))
        return -1;
    if (upload_pop(a, a->g0, g0, gp, false)) return -1;
    ADJ_HIP(hipMemcpy(a->target, target, nspecies * n * 8, hipMemcpyHostToDevice));
    for (int s = 0; s < LBIB_MAX_SPECIES; s++) a->decay[s] = s < nspecies ? decay[s] : 0.0;
    a->have_forward = false;
    return 0;
}

int lbib_adj_set_junctions(lbib_adj* a, int npairs, const int32_t* pairs, double l0) {
    if (!a || npairs < 0 || (npairs > 0 && !pairs)) return -1;
    if (!(l0 >= 0.0)) { a->err = "lbib_adj_set_junctions: l0 must be non-negative"; return -1; }
    const int nm = a->nm;
    std::vector<int32_t> cnt(nm + 1, 0);
    for (int p = 0; p < npairs; p++) {
        const int u = pairs[2 * p], v = pairs[2 * p + 1];
        if (u < 0 || v < 0 || u >= nm || v >= nm || u == v) {
            a->err = "lbib_adj_set_junctions: pair of invalid or equal marker indices";
            return -1;
        }
        cnt[u + 1]++;
        cnt[v + 1]++;
    }
    for (int k = 0; k < nm; k++) cnt[k + 1] += cnt[k];
    std::vector<int32_t> part(std::max(2 * npairs, 1)), cur(cnt.begin(), cnt.end() - 1);
    for (int p = 0; p < npairs; p++) {
        const int u = pairs[2 * p], v = pairs[2 * p + 1];
        part[cur[u]++] = v;
        part[cur[v]++] = u;
    }
    ADJ_HIP(hipStreamSynchronize(a->stream));
    if (a->jofs) (void)hipFree(a->jofs);
    if (a->jpart) (void)hipFree(a->jpart);
    a->jofs = a->jpart = nullptr;
    if (npairs > 0) {
        if (adj_alloc(a, (void**)&a->jofs, (nm + 1) * 4) || adj_alloc(a, (void**)&a->jpart, part.size() * 4))
            return -1;
        ADJ_HIP(hipMemcpy(a->jofs, cnt.data(), (nm + 1) * 4, hipMemcpyHostToDevice));
        ADJ_HIP(hipMemcpy(a->jpart, part.data(), part.size() * 4, hipMemcpyHostToDevice));
    }
    a->l0 = l0;
    a->have_forward = false;
    return 0;
}

int lbib_adj_set_batch_layout(lbib_adj* a, int nrect, const int32_t* rects) {
    if (!a || nrect < 0 || (nrect > 0 && !rects)) return -1;
    if (!a->have_problem) { a->err = "lbib_adj_set_batch_layout: set the problem first"; return -1; }
    const int nx = a->cfg.nx, ny = a->cfg.ny, ntx = (nx + 15) / 16, nty = (ny + 15) / 16;
    std::vector<int32_t> owner((size_t)ntx * nty, -1);
    for (int r = 0; r < nrect; r++) {
        const int x0 = rects[4 * r], y0 = rects[4 * r + 1], w = rects[4 * r + 2], h = rects[4 * r + 3];
        if (x0 < 0 || y0 < 0 || x0 % 16 || y0 % 16 || w < 8 || h < 8 || x0 + w > nx || y0 + h > ny) {
            a->err = "lbib_adj_set_batch_layout: origins must be multiples of 16, sides >= 8, inside the lattice";
            return -1;
        }
        for (int ty = y0 / 16; ty < (y0 + h + 15) / 16; ty++)
            for (int tx = x0 / 16; tx < (x0 + w + 15) / 16; tx++) {
                int32_t& o = owner[(size_t)ty * ntx + tx];
                if (o >= 0) { a->err = "lbib_adj_set_batch_layout: rectangles share a 16x16 tile"; return -1; }
                o = r;
            }
    }
    // every marker starts inside one rectangle, which it keeps: its stencil
    // wraps inside that rectangle from then on
    std::vector<int32_t> mrect(a->xy_host.size() / 2, 0);
    for (size_t k = 0; 2 * k < a->xy_host.size() && nrect > 0; k++) {
        const double X = a->xy_host[2 * k], Y = a->xy_host[2 * k + 1];
        const int r = (X >= 0 && Y >= 0 && X < nx && Y < ny) ? owner[(size_t)((int)Y / 16) * ntx + (int)X / 16] : -1;
        mrect[k] = r;
        const bool in = r >= 0 && X >= rects[4 * r] && X < rects[4 * r] + rects[4 * r + 2] &&
                        Y >= rects[4 * r + 1] && Y < rects[4 * r + 1] + rects[4 * r + 3];
        if (!in) {
            a->err = "lbib_adj_set_batch_layout: every marker must start inside a rectangle";
            return -1;
        }
    }
    ADJ_HIP(hipStreamSynchronize(a->stream));
    if (a->rtile) (void)hipFree(a->rtile);
    if (a->rects) (void)hipFree(a->rects);
    if (a->mrect) (void)hipFree(a->mrect);
    a->rtile = nullptr;
    a->rects = nullptr;
    a->mrect = nullptr;
    a->ntx = ntx;
    if (nrect > 0) {
        if (adj_alloc(a, (void**)&a->rtile, owner.size() * 4) || adj_alloc(a, (void**)&a->rects, (size_t)nrect * 16) ||
            adj_alloc(a, (void**)&a->mrect, std::max<size_t>(mrect.size(), 1) * 4))
            return -1;
        ADJ_HIP(hipMemcpy(a->rtile, owner.data(), owner.size() * 4, hipMemcpyHostToDevice));
        ADJ_HIP(hipMemcpy(a->rects, rects, (size_t)nrect * 16, hipMemcpyHostToDevice));
        if (!mrect.empty())
            ADJ_HIP(hipMemcpy(a->mrect, mrect.data(), mrect.size() * 4, hipMemcpyHostToDevice));
    }
    a->have_forward = false;
    return 0;
}

int lbib_adj_forward_ms(lbib_adj* a, int steps, const double* tau_c, double kappa, double k_junction,
                        const double* K, const double* rates, double* loss, double* c_final) {
    if (!a || !loss || !tau_c || steps < 1) return -1;
    if (!a->have_problem) { a->err = "lbib_adj_forward: no problem set"; return -1; }
    for (int s = 0; s < a->ns; s++)
        if (!(tau_c[s] > 0.5)) { a->err = "lbib_adj_forward: tau_c must exceed 0.5"; return -1; }
    for (int s = 0; s < LBIB_MAX_SPECIES; s++) a->tau_c[s] = s < a->ns ? tau_c[s] : 0.8;
    for (int s = 0; s < LBIB_MAX_SPECIES; s++)
        for (int j = 0; j < LBIB_MAX_SPECIES; j++)
            a->Km[s * LBIB_MAX_SPECIES + j] = (K && s < a->ns && j < a->ns) ? K[s * a->ns + j] : 0.0;
    a->kappa = kappa;
    a->kj = k_junction;
    for (int c = 0; c <= a->ncell; c++) a->rates[c] = rates && c > 0 ? rates[c] : 0.0;
    ADJ_HIP(hipMemcpy(a->src, a->rates.data(), (a->ncell + 1) * 8, hipMemcpyHostToDevice));
    const int r = a->prec == LBIB_FP64 ? forward_t<double>(a, steps, loss, c_final)
                                       : forward_t<float>(a, steps, loss, c_final);
    a->have_forward = r == 0;
    return r;
}

int lbib_adj_gradient_ms(lbib_adj* a, double* d_tau_c, double* d_kappa, double* d_k_junction, double* d_K,
                         double* d_rates) {
    if (!a) return -1;
    if (!a->have_forward) { a->err = "lbib_adj_gradient: run lbib_adj_forward first"; return -1; }
    const int r = a->prec == LBIB_FP64 ? gradient_t<double>(a, d_tau_c, d_kappa, d_k_junction, d_K, d_rates)
                                       : gradient_t<float>(a, d_tau_c, d_kappa, d_k_junction, d_K, d_rates);
    // the loss adjoint in lg[0] was consumed: a new forward is needed
    a->have_forward = false;
    return r;
}

int lbib_adj_forward(lbib_adj* a, int steps, double tau_c, double kappa, const double* rates, double* loss,
                     double* c_final) {
    if (!a) return -1;
    if (a->ns != 1) { a->err = "lbib_adj_forward: several species; use lbib_adj_forward_ms"; return -1; }
    return lbib_adj_forward_ms(a, steps, &tau_c, kappa, 0.0, nullptr, rates, loss, c_final);
}

int lbib_adj_gradient(lbib_adj* a, double* d_tau_c, double* d_kappa, double* d_rates) {
    if (!a) return -1;
    if (a->ns != 1) { a->err = "lbib_adj_gradient: several species; use lbib_adj_gradient_ms"; return -1; }
    return lbib_adj_gradient_ms(a, d_tau_c, d_kappa, nullptr, nullptr, d_rates);
}

int lbib_adj_set_storage(lbib_adj* a, int storage) {
    if (!a) return -1;
    if (a->have_problem) { a->err = "lbib_adj_set_storage: choose the storage before lbib_adj_set_problem"; return -1; }
    const bool same = storage == a->prec;
    const bool half = (storage == LBIB_FP16_DEV || storage == LBIB_BF16_DEV) && a->prec == LBIB_FP32;
    if (!same && !half) {
        a->err = "lbib_adj_set_storage: the arithmetic precision, or FP16 / BF16 deviations with FP32 arithmetic";
        return -1;
    }
    ADJ_HIP(hipStreamSynchronize(a->stream));
    a->sprec = storage;
    a->sb = same ? a->tb : 2;
    a->have_forward = false;
    return alloc_states(a);
}

int lbib_adj_set_checkpointing(lbib_adj* a, int scheme, int param) {
    if (!a) return -1;
    if ((scheme != LBIB_ADJ_CHECKPOINT_EVERY && scheme != LBIB_ADJ_REVOLVE) || param < 1) {
        a->err = "lbib_adj_set_checkpointing: LBIB_ADJ_CHECKPOINT_EVERY with K >= 1 steps, or LBIB_ADJ_REVOLVE "
                 "with >= 1 snapshot slots";
        return -1;
    }
    ADJ_HIP(hipStreamSynchronize(a->stream));
    adj_free(a->seg);
    adj_free(a->ck);
    a->nck = 0;
    a->scheme = scheme == LBIB_ADJ_REVOLVE ? 1 : 0;
    if (a->scheme == 1) {
        a->snaps = param;
    } else {
        a->K = param;
        if (adj_alloc(a, &a->seg, (size_t)a->K * (state_bytes(a) + extra_bytes(a)))) return -1;
    }
    a->have_forward = false;
    return 0;
}

int lbib_adj_stats(const lbib_adj* a, int64_t* forward_steps, int64_t* gradient_steps, int64_t* store_bytes) {
    if (!a) return -1;
    if (forward_steps) *forward_steps = a->fwd_steps_run;
    if (gradient_steps) *gradient_steps = a->grad_steps_run;
    if (store_bytes) {
        const size_t ck = (size_t)(a->scheme == 1 ? std::max(a->snaps, 1) : a->steps / a->K + 1) * state_bytes(a);
        const size_t seg = a->scheme == 0 ? (size_t)a->K * (state_bytes(a) + extra_bytes(a)) : 0;
        *store_bytes = (int64_t)(ck + seg);
    }
    return 0;
}

int lbib_adj_timing(lbib_adj* a, double* forward_ms, double* gradient_ms) {
    if (!a) return -1;
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (gradient_ms) *gradient_ms = a->grad_ms;
    return 0;
}

}  // extern "C"
