// lbib_mixed.h -- shared pieces of the mixed-precision lattice kernels (FP32
// arithmetic on populations stored in FP32, FP16/BF16 deviations, or FP32
// moments with FP16 remainders): lattice constants, the kernel argument
// block, and the deviation-form BGK and D2Q5 collisions.
#pragma once
#include "lbib_device_math.h"

namespace lbib {

// lattice weights and velocities in natural order T,E,N,W,S,NE,NW,SW,SE
__device__ __forceinline__ float wq(int q) { return q == 0 ? 4.0f / 9.0f : (q < 5 ? 1.0f / 9.0f : 1.0f / 36.0f); }
__device__ __forceinline__ int exq(int q) {
    const int e[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    return e[q];
}
__device__ __forceinline__ int eyq(int q) {
    const int e[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    return e[q];
}

struct SpeciesF {
    int kind;
    float tauI;
    float production, decay;
    uint32_t switchoff;
    int prod_domain;
    float source_inside;
    float krow[LBIB_MAX_SPECIES];
};

struct MixedArgs {
    int nx, ny;
    long n;
    const void* fs;
    void* fd;
    const void* gs;
    void* gd;
    float* u;            // velocity pairs
    const float2* F;     // force pairs (FP32)
    const uint8_t* ftile;
    const uint32_t* tag;
    const double* mtab;
    int ntab;
    double* msite;
    const double* bc_v;
    const int32_t* bc_desc;
    const int32_t* bcd_off;
    const int32_t* bcd_sp;
    int ntx;
    float tauI, tau;
    uint32_t it;
    int box_outlet;
    const DevStatus* st;
    const int32_t* rtile;   // batched atlas: rectangle per 16x16 tile (null: one domain)
    const RectDev* rects;
    SpeciesF sp[LBIB_MAX_SPECIES];
};

// Deviation-form BGK with Luo (FM 0) or Guo (FM 1 as written, FM 2 textbook)
// forcing; d[] in, d[] out, post-collision velocity out.  rho - 1 enters the
// equilibrium as the deviation sum itself, never as (1 + sum) - 1.
template <int FM>
__device__ __forceinline__ void collide_dev_f32(float d[9], float m, float Fx, float Fy, float tauI, float tau,
                                                float gx, float gy, float& uxo, float& uyo) {
    float sd = 0.0f;
#pragma unroll
    for (int q = 0; q < 9; q++) sd += d[q];
    // the density deviation rho - 1 (with the mass source) is kept apart from
    // the 1: a source of 1e-5 added to rho itself would be rounded to the
    // spacing of FP32 at 1 (1.2e-7), to the spacing at |rho - 1| here
    float dr = sd;
    if (isfinite(m)) dr += m;
    const float rho0 = 1.0f + sd;
    const float rho = 1.0f + dr;
    const float rhoI = 1.0f / rho;
    float jx = d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW]);
    float jy = d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE]);
    if (FM != 0) { jx += 0.5f * Fx; jy += 0.5f * Fy; }
    const float ux = jx * rhoI, uy = jy * rhoI;
    const float u2 = 1.5f * (ux * ux + uy * uy);
    const float G = (FM != 0) ? (1.0f - 0.5f / tau) : 0.0f;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const float w = wq(q);
        const float ex = (float)exq(q), ey = (float)eyq(q);
        const float eu = ex * ux + ey * uy;
        const float deq = w * dr + w * rho * (3.0f * eu + 4.5f * eu * eu - u2);
        float fc;
        if (FM == 0) {
            fc = 3.0f * w * rho0 * (ex * Fx + ey * Fy);
        } else {
            // w (1 - 1/(2 tau)) [3 (e - u) + 9 (e.u) e] . F with the stored velocity
            const float eg = ex * gx + ey * gy;
            float cx = 3.0f * (ex - gx) + 9.0f * eg * ex;
            float cy = 3.0f * (ey - gy) + 9.0f * eg * ey;
            if (FM == 1 && q == DNE) cy = 9.0f * gx - 3.0f * gy + 3.0f;   // GuoZhengShi2002ForceModel.cpp NE y-term
            fc = G * w * (cx * Fx + cy * Fy);
        }
        d[q] = d[q] - (d[q] - deq) * tauI + fc;
    }
    float px = d[DE] + d[DNE] + d[DSE] - (d[DNW] + d[DW] + d[DSW]);
    float py = d[DNE] + d[DN] + d[DNW] - (d[DSW] + d[DS] + d[DSE]);
    if (FM != 0) { px += 0.5f * Fx; py += 0.5f * Fy; }
    uxo = px * rhoI;
    uyo = py * rhoI;
}

template <int NS>
__device__ __forceinline__ float reaction_f32(const SpeciesF& p, uint32_t dom, uint32_t it, float C, const float* cj,
                                              const float* cseq) {
    if (p.kind == LBIB_SPECIES_SCHNAKENBERG_R || p.kind == LBIB_SPECIES_SCHNAKENBERG_L) {
        if (dom == 0u) return 0.0f;
        const float Cp = cseq[p.prod_domain];
        if (p.kind == LBIB_SPECIES_SCHNAKENBERG_R) return p.production * (p.source_inside - C + C * C * Cp);
        return p.production * (p.source_inside - Cp * Cp * C);
    }
    if (p.kind == LBIB_SPECIES_SIGNAL) {
        const bool prod = (p.prod_domain < 0) ? (dom != 0u) : (dom == (uint32_t)p.prod_domain);
        if (prod && it < p.switchoff) return p.production;
        return -p.decay * C;
    }
    if (p.kind == LBIB_SPECIES_LINEAR) {
        float r = 0.0f;
#pragma unroll
        for (int j = 0; j < NS; j++) r += p.krow[j] * cj[j];
        if (dom != 0u) r += p.source_inside;
        return r;
    }
    return 0.0f;
}

template <int NS>
__device__ __forceinline__ void species_collide_f32(float g[5], float C, float ux, float uy, const SpeciesF& p,
                                                    float tauI, uint32_t dom, uint32_t it, const float* cj,
                                                    const float* cseq) {
    const float w0 = C / 3.0f, w1 = C / 6.0f;
    float eq[5];
    eq[DT] = w0;
    eq[DE] = w1 * (1.0f + 3.0f * ux);
    eq[DN] = w1 * (1.0f + 3.0f * uy);
    eq[DW] = w1 * (1.0f - 3.0f * ux);
    eq[DS] = w1 * (1.0f - 3.0f * uy);
    float R0 = 0.0f, R1 = 0.0f;
    if (p.kind != LBIB_SPECIES_PLAIN) {
        float R = reaction_f32<NS>(p, dom, it, C, cj, cseq);
        if (p.kind == LBIB_SPECIES_SCHNAKENBERG_R || p.kind == LBIB_SPECIES_SCHNAKENBERG_L) R *= p.decay;
        R0 = R / 3.0f;
        R1 = R / 6.0f;
    }
    g[DT] = g[DT] - (g[DT] - eq[DT]) * tauI + R0;
#pragma unroll
    for (int q = 1; q < 5; q++) g[q] = g[q] - (g[q] - eq[q]) * tauI + R1;
}

inline MixedArgs make_mixed_args(lbib_ctx* c, uint32_t it, bool use_ftile) {
    MixedArgs a;
    a.nx = c->nx; a.ny = c->ny; a.n = c->n;
    a.fs = c->f[c->cur];
    a.fd = c->f[1 - c->cur];
    a.gs = c->g[c->cur];
    a.gd = c->g[1 - c->cur];
    a.u = reinterpret_cast<float*>(c->u);
    a.F = reinterpret_cast<const float2*>(c->Fx);
    a.ftile = use_ftile ? c->tact : nullptr;
    a.tag = c->tag;
    a.mtab = c->mtab_live ? c->mtab : nullptr;
    a.ntab = c->ndom;
    a.msite = c->msite_live ? c->msite : nullptr;
    a.bc_v = c->bc_v; a.bc_desc = c->bc_desc; a.bcd_off = c->bcd_off; a.bcd_sp = c->bcd_sp;
    a.ntx = c->stx;
    a.tauI = (float)(1.0 / c->cfg.tau_fluid);
    a.tau = (float)c->cfg.tau_fluid;
    a.it = it;
    a.box_outlet = c->cfg.box_outlet;
    a.st = c->st;
    a.rtile = c->nrect ? c->rtile : nullptr;
    a.rects = c->rects;
    for (int s = 0; s < LBIB_MAX_SPECIES; s++) {
        const SpeciesDev& p = c->sp[s];
        SpeciesF& q = a.sp[s];
        q.kind = p.kind;
        q.tauI = (float)p.tauI;
        q.production = (float)p.production;
        q.decay = (float)p.decay;
        q.switchoff = p.switchoff;
        q.prod_domain = p.prod_domain;
        q.source_inside = (float)p.source_inside;
        for (int j = 0; j < LBIB_MAX_SPECIES; j++) q.krow[j] = (float)p.krow[j];
    }
    return a;
}

// LBIB_FP16_REM (lbib_rem.hip): lattice launch, and FP64 natural-order
// fields to and from the moments-and-remainders store.
int launch_rem_shape(lbib_ctx* c, const MixedArgs& a, int variant, int wg);
int launch_rem_pack(lbib_ctx* c, const double* src_dev, int field);
int launch_rem_unpack(lbib_ctx* c, double* dst_dev, int field);

}  // namespace lbib
