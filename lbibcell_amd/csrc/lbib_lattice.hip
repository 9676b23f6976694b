// lbib_lattice.hip -- fused D2Q9 + D2Q5 pull-stream / collide kernel (FP64).
//
// One thread per site, 256-wide row tiles.  Each thread pulls its 9 + 5*ns
// populations from the neighbours (periodic, GeometryHandler.cpp:520-589),
// applies the no-flux cut-link override (BoundarySolverNoFluxD2Q5.cpp:37-113),
// collides the fluid and then every species with the fresh velocity
// (SimulationRunner.cpp:251-258), applies the MassSolverBoxOutlet rescale on
// x = 0 and y = 0 (MassSolverBoxOutlet.cpp:42-72) and writes post-collision
// natural-order values.  Equivalent to the reference's swap streaming
// (FluidSolver.cpp:185-194): after the swaps slot d of site x holds
// f*_d(x - e_d), which is what the pull reads.
//
// Roofline: HBM bound.  Algorithmic bytes per site = 2*9*8 + ns*2*5*8.
#include "lbib_device_math.h"

namespace lbib {

struct LatticeArgs {
    int nx, ny;
    long n;
    const double* __restrict__ fs;
    double* __restrict__ fd;
    const double* __restrict__ gs;
    double* __restrict__ gd;
    double* u;           // velocity pairs (ux, uy)
    const double* Fx;
    const double* Fy;
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
    double tauI, tau;
    uint32_t it;
    int box_outlet;
    const DevStatus* st;   // the step is skipped while the status is stopped
    const int32_t* rtile;  // batched atlas: rectangle per 16x16 tile (null: one domain)
    const RectDev* rects;
    SpeciesDev sp[LBIB_MAX_SPECIES];
};

template <int NS, int FM>
__global__ __launch_bounds__(256) void k_stream_collide_f64(const LatticeArgs a) {
    if (step_stopped(a.st)) return;
    const uint32_t wg = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int x = (int)(wg % gridDim.x) * (int)blockDim.x + threadIdx.x;
    const int y = (int)(wg / gridDim.x);
    if (x >= a.nx) return;
    Frame fr;
    if (!site_frame(a.rtile, a.rects, a.ntx, a.nx, a.ny, x, y, fr)) return;
    const long n = a.n;
    const int xm = (x == fr.x0) ? fr.x0 + fr.w - 1 : x - 1;
    const int xp = (x == fr.x0 + fr.w - 1) ? fr.x0 : x + 1;
    const long r0 = (long)y * a.nx;
    const long rS = (long)((y == fr.y0) ? fr.y0 + fr.h - 1 : y - 1) * a.nx;
    const long rN = (long)((y == fr.y0 + fr.h - 1) ? fr.y0 : y + 1) * a.nx;
    const long i = r0 + x;
    const double tauI = fr.r < 0 ? a.tauI : a.rects[fr.r].tauI;
    const double tau = fr.r < 0 ? a.tau : a.rects[fr.r].tau;

    // every population is read once per step: non-temporal loads
    double f[9];
    f[DT] = __builtin_nontemporal_load(&a.fs[i]);
    f[DE] = __builtin_nontemporal_load(&a.fs[1 * n + r0 + xm]);
    f[DN] = __builtin_nontemporal_load(&a.fs[2 * n + rS + x]);
    f[DW] = __builtin_nontemporal_load(&a.fs[3 * n + r0 + xp]);
    f[DS] = __builtin_nontemporal_load(&a.fs[4 * n + rN + x]);
    f[DNE] = __builtin_nontemporal_load(&a.fs[5 * n + rS + xm]);
    f[DNW] = __builtin_nontemporal_load(&a.fs[6 * n + rS + xp]);
    f[DSW] = __builtin_nontemporal_load(&a.fs[7 * n + rN + xp]);
    f[DSE] = __builtin_nontemporal_load(&a.fs[8 * n + rN + xm]);

    const uint32_t tg = a.tag[i];
    const uint32_t dom = tg & DOM_MASK;
    double Fx = 0.0, Fy = 0.0;
    if (a.ftile == nullptr || a.ftile[(long)(y >> 4) * a.ntx + (x >> 4)]) {
        Fx = a.Fx[i];
        Fy = a.Fy[i];
    }
    double m = 0.0;
    if (a.msite) { m = a.msite[i]; a.msite[i] = 0.0; }
    else if (a.mtab) m = (dom < (uint32_t)a.ntab) ? a.mtab[dom] : 0.0;
    // stored velocity_ of the previous collide: read only where it is used
    // (Guo forcing everywhere, the no-flux BC at boundary-node sites)
    double pu = 0.0, pv = 0.0;
    if (FM != 0 || (NS > 0 && (tg >> BC_SHIFT))) {
        pu = a.u[2 * i];
        pv = a.u[2 * i + 1];
    }

    double g[NS > 0 ? NS : 1][5];
    const long sn = 5 * n;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double* gs = a.gs + s * sn;
        g[s][DT] = __builtin_nontemporal_load(&gs[i]);
        g[s][DE] = __builtin_nontemporal_load(&gs[1 * n + r0 + xm]);
        g[s][DN] = __builtin_nontemporal_load(&gs[2 * n + rS + x]);
        g[s][DW] = __builtin_nontemporal_load(&gs[3 * n + r0 + xp]);
        g[s][DS] = __builtin_nontemporal_load(&gs[4 * n + rN + x]);
    }
    if (NS > 0 && (tg >> BC_SHIFT)) {
        // Cut links carrying BoundarySolverNoFluxD2Q5: slot o holds a boundary
        // node whose direction to this site is d = inverse(o); the population
        // arriving in direction d is replaced (preAdvect computes it from the
        // first listed species only, postAdvect swaps it into every listed
        // species in turn).
        const int inv[5] = {DT, DW, DS, DE, DN};
        for (int o = 1; o <= 4; o++) {
            if (!((tg >> (BC_SHIFT + o - 1)) & 1u)) continue;
            const int d = inv[o];
            const long e = (long)(o - 1) * n + i;
            const int desc = a.bc_desc[e];
            const int b0 = a.bcd_off[desc], b1 = a.bcd_off[desc + 1];
            if (b1 <= b0) continue;
            const int s1 = a.bcd_sp[b0];
            const double* gs = a.gs + s1 * sn;
            const double Cp = c_swapped(gs[i], gs[1 * n + i], gs[2 * n + i], gs[3 * n + i], gs[4 * n + i]);
            const double Cf = Cp / 6.0;
            const double w1 = Cp / 6.0;
            const double bvx = a.bc_v[2 * e], bvy = a.bc_v[2 * e + 1];
            double feq = 0.0, eq = 0.0;
            if (d == DE) { feq = Cf * (1 + 3 * bvx); eq = w1 * (1.0 + pu * 3.0); }
            else if (d == DN) { feq = Cf * (1 + 3 * bvy); eq = w1 * (1.0 + pv * 3.0); }
            else if (d == DW) { feq = Cf * (1 - 3 * bvx); eq = w1 * (1.0 + (-pu) * 3.0); }
            else { feq = Cf * (1 - 3 * bvy); eq = w1 * (1.0 + (-pv) * 3.0); }
            const double fneq = gs[(long)d * n + i] - eq;
            double carry = feq + fneq;
            for (int k = b0; k < b1; k++) {
                const int sk = a.bcd_sp[k];
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    if (s == sk) { const double t = g[s][d]; g[s][d] = carry; carry = t; }
                }
            }
        }
    }

    double uxn, uyn;
    fluid_collide<FM>(f, m, Fx, Fy, tauI, tau, pu, pv, uxn, uyn);

    if (NS > 0) {
        double cj[NS > 0 ? NS : 1], cseq[NS > 0 ? NS : 1];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            double c = 0.0;
#pragma unroll
            for (int q = 0; q < 5; q++) c += g[s][q];
            cj[s] = c;
            cseq[s] = c;
        }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const double tI = fr.r < 0 ? a.sp[s].tauI : a.rects[fr.r].tauIs[s];
            species_collide<NS>(g[s], cj[s], uxn, uyn, a.sp[s], tI, dom, iter_of(a.it, a.st), cj, cseq);
            // later species read this one's C from its post-collision slots
            cseq[s] = c_swapped(g[s][DT], g[s][DE], g[s][DN], g[s][DW], g[s][DS]);
        }
    }

    if (a.box_outlet && (x == fr.x0 || y == fr.y0)) {
        // x = 0 column first, then the y = 0 row; (0,0) is rescaled twice.
        int passes = (x == fr.x0) + (y == fr.y0);
        for (int p = 0; p < passes; p++) {
            const double sf = 1.0 / rho_swapped(f);
#pragma unroll
            for (int q = 0; q < 9; q++) f[q] *= sf;
#pragma unroll
            for (int s = 0; s < NS; s++)
#pragma unroll
                for (int q = 0; q < 5; q++) g[s][q] *= sf;
        }
    }

    // streaming stores: the next pass reads these from HBM, not from cache
#pragma unroll
    for (int q = 0; q < 9; q++) __builtin_nontemporal_store(f[q], &a.fd[(long)q * n + i]);
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) __builtin_nontemporal_store(g[s][q], &a.gd[s * sn + (long)q * n + i]);
    *reinterpret_cast<double2*>(a.u + 2 * i) = make_double2(uxn, uyn);
}

// FluidSolver::initSolver / initWithVelocity and the species initSolver.
template <int NS, int FM>
__global__ void k_init_f64(const LatticeArgs a, const double* u0) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const long n = a.n;
    // sites outside every rectangle of a batched atlas take the defaults
    Frame fr;
    if (!site_frame(a.rtile, a.rects, a.ntx, a.nx, a.ny, (int)(i % a.nx), (int)(i / a.nx), fr)) fr.r = -1;
    double f[9];
    double uxn, uyn;
    if (u0 == nullptr) {
        // FluidSolver.cpp:50-60: distributions_[T] = 1, collide with F = 0, m = 0
#pragma unroll
        for (int q = 0; q < 9; q++) f[q] = 0.0;
        f[DT] = 1.0;
        const double tauI = fr.r < 0 ? a.tauI : a.rects[fr.r].tauI;
        const double tau = fr.r < 0 ? a.tau : a.rects[fr.r].tau;
        fluid_collide<FM>(f, 0.0, 0.0, 0.0, tauI, tau, 0.0, 0.0, uxn, uyn);
    } else {
        // FluidSolver.cpp:62-101
        const double u = u0[2 * i], v = u0[2 * i + 1];
        const double rho = 1.0, u2 = u * u, v2 = v * v;
        const double W1rho = kW1 * rho, W2rho = kW2 * rho, W3 = 9.0 / 2.0;
        const double lastTerm = 3.0 / 2.0 * (u2 + v2);
        f[DT] = kW0 * rho * (1.0 - lastTerm);
        f[DE] = W1rho * (1.0 + 3.0 * u + W3 * u2 - lastTerm);
        f[DNE] = W2rho * (1.0 + 3.0 * (u + v) + W3 * (u + v) * (u + v) - lastTerm);
        f[DN] = W1rho * (1.0 + 3.0 * v + W3 * v2 - lastTerm);
        f[DNW] = W2rho * (1.0 + 3.0 * (-u + v) + W3 * (-u + v) * (-u + v) - lastTerm);
        f[DW] = W1rho * (1.0 + 3.0 * (-u) + W3 * u2 - lastTerm);
        f[DSW] = W2rho * (1.0 + 3.0 * (-u - v) + W3 * (-u - v) * (-u - v) - lastTerm);
        f[DS] = W1rho * (1.0 + 3.0 * (-v) + W3 * v2 - lastTerm);
        f[DSE] = W2rho * (1.0 + 3.0 * (u - v) + W3 * (u - v) * (u - v) - lastTerm);
        uxn = u;
        uyn = v;
    }
#pragma unroll
    for (int q = 0; q < 9; q++) a.fd[(long)q * n + i] = f[q];
    *reinterpret_cast<double2*>(a.u + 2 * i) = make_double2(uxn, uyn);
    const uint32_t dom = a.tag[i] & DOM_MASK;
    const long sn = 5 * n;
    for (int s = 0; s < NS; s++) {
        double g[5];
        if (a.sp[s].kind != LBIB_SPECIES_PLAIN) {
            const double v = (dom == 0u) ? 0.0 : a.sp[s].init_inside / 5.0;
#pragma unroll
            for (int q = 0; q < 5; q++) g[q] = v;
        } else {
            // CDESolverD2Q5.cpp:37-40
            g[0] = 1.0; g[1] = 0.0; g[2] = 0.0; g[3] = 0.0; g[4] = 0.0;
            const double tI = fr.r < 0 ? a.sp[s].tauI : a.rects[fr.r].tauIs[s];
            species_collide<NS>(g, 1.0, uxn, uyn, a.sp[s], tI, dom, a.it, nullptr, nullptr);
        }
#pragma unroll
        for (int q = 0; q < 5; q++) a.gd[s * sn + (long)q * n + i] = g[q];
    }
}

static LatticeArgs make_args(lbib_ctx* c, uint32_t it, bool use_ftile) {
    LatticeArgs a;
    a.nx = c->nx; a.ny = c->ny; a.n = c->n;
    a.fs = (const double*)c->f[c->cur];
    a.fd = (double*)c->f[1 - c->cur];
    a.gs = (const double*)c->g[c->cur];
    a.gd = (double*)c->g[1 - c->cur];
    a.u = c->u;
    a.Fx = c->Fx; a.Fy = c->Fy;
    a.ftile = use_ftile ? c->tact : nullptr;
    a.tag = c->tag;
    a.mtab = c->mtab_live ? c->mtab : nullptr;
    a.ntab = c->ndom;
    a.msite = c->msite_live ? c->msite : nullptr;
    a.bc_v = c->bc_v; a.bc_desc = c->bc_desc; a.bcd_off = c->bcd_off; a.bcd_sp = c->bcd_sp;
    a.ntx = c->stx;
    a.tauI = 1.0 / c->cfg.tau_fluid;
    a.tau = c->cfg.tau_fluid;
    a.it = it;
    a.box_outlet = c->cfg.box_outlet;
    a.st = c->st;
    a.rtile = c->nrect ? c->rtile : nullptr;
    a.rects = c->rects;
    for (int s = 0; s < LBIB_MAX_SPECIES; s++) a.sp[s] = c->sp[s];
    return a;
}

template <int FM>
static void dispatch_sc(lbib_ctx* c, const LatticeArgs& a, int wg) {
    dim3 grid((c->nx + wg - 1) / wg, c->ny);
    switch (c->ns) {
        case 0: k_stream_collide_f64<0, FM><<<grid, wg, 0, c->stream>>>(a); break;
        case 1: k_stream_collide_f64<1, FM><<<grid, wg, 0, c->stream>>>(a); break;
        case 2: k_stream_collide_f64<2, FM><<<grid, wg, 0, c->stream>>>(a); break;
        case 3: k_stream_collide_f64<3, FM><<<grid, wg, 0, c->stream>>>(a); break;
        default: k_stream_collide_f64<4, FM><<<grid, wg, 0, c->stream>>>(a); break;
    }
}

static int launch_f64_shape(lbib_ctx* c, const LatticeArgs& a, int wg) {
    if (c->cfg.force_model == LBIB_FORCE_LUO) dispatch_sc<0>(c, a, wg);
    else if (c->cfg.force_model == LBIB_FORCE_GUO_AS_WRITTEN) dispatch_sc<1>(c, a, wg);
    else dispatch_sc<2>(c, a, wg);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_stream_collide_f64(lbib_ctx* c, uint32_t it, bool use_ftile) {
    LatticeArgs a = make_args(c, it, use_ftile);
    if (launch_f64_shape(c, a, lattice_shape(c).wg)) return -1;
    c->cur = 1 - c->cur;
    c->mtab_live = false;
    c->msite_live = false;
    return 0;
}

template <int FM>
static void dispatch_init(lbib_ctx* c, const LatticeArgs& a, const double* u0) {
    const int bs = 256;
    const int nb = (int)((c->n + bs - 1) / bs);
    switch (c->ns) {
        case 0: k_init_f64<0, FM><<<nb, bs, 0, c->stream>>>(a, u0); break;
        case 1: k_init_f64<1, FM><<<nb, bs, 0, c->stream>>>(a, u0); break;
        case 2: k_init_f64<2, FM><<<nb, bs, 0, c->stream>>>(a, u0); break;
        case 3: k_init_f64<3, FM><<<nb, bs, 0, c->stream>>>(a, u0); break;
        default: k_init_f64<4, FM><<<nb, bs, 0, c->stream>>>(a, u0); break;
    }
}

int launch_init_f64(lbib_ctx* c, const double* u0_dev) {
    // init writes the *current* buffers
    return launch_init_f64_into(c, u0_dev, (double*)c->f[c->cur], (double*)c->g[c->cur], c->u);
}

int launch_init_f64_into(lbib_ctx* c, const double* u0_dev, double* f, double* g, double* u) {
    LatticeArgs a = make_args(c, 0, false);
    a.fd = f;
    a.gd = g;
    a.u = u;
    if (c->cfg.force_model == LBIB_FORCE_LUO) dispatch_init<0>(c, a, u0_dev);
    else if (c->cfg.force_model == LBIB_FORCE_GUO_AS_WRITTEN) dispatch_init<1>(c, a, u0_dev);
    else dispatch_init<2>(c, a, u0_dev);
    LBIB_HIP(hipGetLastError());
    return 0;
}

// Plain streaming copy on the context stream: the attainable-bandwidth
// yardstick for the lattice kernel (same nontemporal store path) and the
// scale for hardware byte counters.
typedef double probe_vec __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy_probe(const probe_vec* __restrict__ src, probe_vec* __restrict__ dst,
                                                    long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(src[i], &dst[i]);
}

int launch_copy_probe(lbib_ctx* c, const void* src, void* dst, long nbytes) {
    const long n = nbytes / (long)sizeof(probe_vec);
    const long nb = (n + 255) / 256;
    if (n <= 0 || nb > 0x7fffffffL) {
        c->err = "copy probe size out of range";
        return -1;
    }
    k_copy_probe<<<(unsigned)nb, 256, 0, c->stream>>>((const probe_vec*)src, (probe_vec*)dst, n);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int autotune_f64(lbib_ctx* c, int reps, std::vector<TunedShape>& out) {
    TuneScratch sc;
    if (sc.init(c, 2 * (size_t)c->n * sizeof(double))) return -1;
    // the active-tile flags are cleared after every step: the timed launches
    // read the force field at every site (the most a step reads)
    LatticeArgs a = make_args(c, 0, false);
    a.u = (double*)sc.u;
    a.msite = nullptr;
    a.mtab = nullptr;
    a.st = sc.st;
    for (int wg : {64, 128, 256}) {
        double ms = 0.0;
        if (sc.time(c, reps, [&] { return launch_f64_shape(c, a, wg); }, ms)) return -1;
        out.push_back({{0, wg, 1}, ms});
    }
    return 0;
}

}  // namespace lbib
