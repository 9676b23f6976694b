// lbib_ib.hip -- immersed-boundary kernels: marker forces, ordered force
// spreading, velocity interpolation + marker move.
//
// Stencil.  GeometryHandler::connectGeometryNodesToPhysicalNodes
// (GeometryHandler.cpp:608-644) links a marker at (X,Y) to the 4x4 sites
// x in [ceil(X-2), floor(X+2)], y likewise, slot (4x+y) % 16, and re-links
// only when the marker leaves the half-open cell of its previous position
// (GeometryNode.cpp:256-261).  For non-integer X (integers are removed by
// GeometryNode::perturbatePosition) that set is x = floor(X)-1 .. floor(X)+2,
// so the kernels derive it from the current position instead of storing it.
//
// Ordered spread.  The reference adds delta*F into each site in marker-id
// order (SimulationRunner.cpp:449-452).  To reproduce those FP64 sums exactly
// the spread is a 4-kernel bucket sort: count contributions per site (first
// toucher lists the site), allocate per-site pool ranges, scatter
// (marker, contribution) records, then per site sort by marker and add in
// that order.  Atomics only count; no floating-point atomics.
#include <algorithm>
#include <climits>
#include "lbib_device_math.h"

namespace lbib {

// ForceSolver::calculateForce (ForceSolver.cpp:42-85) with the force structs of
// ForceStructs.hpp:49-588; per marker, structs in map/push order.
// Also the first kernel of the lattice stage: opens a pending iteration
// (lbib_step_begin) and clears the spread tiles.
__global__ void k_marker_force(const double* __restrict__ mx, const double* __restrict__ my,
                               const int32_t* __restrict__ off, const int32_t* __restrict__ type,
                               const int32_t* __restrict__ partner, const double* __restrict__ prm,
                               double* __restrict__ mfx, double* __restrict__ mfy, int ntiles, uint32_t* tcnt,
                               uint8_t* tact, int begin, uint32_t begin_it, DevStatus* st) {
    if (begin && blockIdx.x == 0 && threadIdx.x == 0) step_begin_dev(st, begin_it);
    if (step_stopped(st)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->ntouched = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        tcnt[t] = 0;
        tact[t] = 0;
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int)st->nm) return;
    const double x1 = mx[k], y1 = my[k];
    double fx = 0.0, fy = 0.0;
    for (int j = off[k]; j < off[k + 1]; j++) {
        const double* p = prm + 4 * (long)j;
        const int t = type[j];
        if (t == 0 || t == 7 || t == 5 || t == 6) {
            const int q = partner[j];
            const double x2 = mx[q], y2 = my[q];
            const double dist = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
            if (t == 5 || t == 6) {
                fx += p[0] * (x2 - x1) / dist;
                fy += p[0] * (y2 - y1) / dist;
            } else {
                const double f = p[0] * (sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1)) - p[1]);
                fx += f * (x2 - x1) / dist;
                fy += f * (y2 - y1) / dist;
            }
        } else if (t == 1) {
            const double d0 = sqrt((p[0] - x1) * (p[0] - x1) + (p[1] - y1) * (p[1] - y1));
            const double f = p[2] * (d0 - p[3]);
            const double dist = sqrt((p[0] - x1) * (p[0] - x1) + (p[1] - y1) * (p[1] - y1));
            if (dist != 0.0) {
                fx += f * (p[0] - x1) / dist;
                fy += f * (p[1] - y1) / dist;
            }
        } else if (t == 2) {
            fx += p[0];
            fy += p[1];
        } else if (t == 3) {
            const double f = p[1] * (sqrt((x1 - x1) * (x1 - x1) + (p[0] - y1) * (p[0] - y1)) - p[2]);
            fx += 0;
            fy += (y1 < p[0]) ? f : -f;
        } else if (t == 4) {
            const double f = p[1] * (sqrt((p[0] - x1) * (p[0] - x1) + (y1 - y1) * (y1 - y1)) - p[2]);
            fx += (x1 < p[0]) ? f : -f;
            fy += 0;
        }
    }
    mfx[k] = fx;
    mfy[k] = fy;
}

// ---- ordered spread over 16x16 site tiles --------------------------------
// Markers are binned by the tile of floor(X), floor(Y).  A tile's sites can be
// reached only by markers binned in the 3x3 tiles around it (the stencil
// reaches 2 sites).  One workgroup per tile gathers the markers whose stencil
// overlaps it, ranks them by index (= GeometryNode id order; indices are
// distinct, so a marker's rank is the number of smaller ones), evaluates each
// marker's eight delta factors once, and every lane adds delta*F for its site
// in rank order, starting from 0.0 like FluidSolver::resetForce.  Tiles with
// no marker nearby are flagged inactive and the lattice kernel reads no force.
constexpr int ST = 16;            // spread tile edge
constexpr int SP_IDX = 1024;      // markers gathered per tile
constexpr int SP_CH = 128;        // marker data staged per pass

__global__ void k_tile_count(int ntx, int nty, const double* __restrict__ mx, const double* __restrict__ my,
                             uint32_t* tcnt, uint32_t* tslot, uint8_t* tact, const DevStatus* st) {
    if (step_stopped(st)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = k < (int)st->nm;
    int fx = 0, fy = 0, home = -1;
    if (live) {
        fx = (int)floor(mx[k]);
        fy = (int)floor(my[k]);
        home = (fy / ST) * ntx + fx / ST;
    }
    // consecutive markers (one membrane) mostly share a home tile: the head
    // lane of each run of equal tiles in the wavefront takes one atomic
    const int lane = threadIdx.x & 63;
    const int hprev = __shfl_up(home, 1, 64);
    const bool head = live && (lane == 0 || hprev != home);
    const unsigned long long heads = __ballot(head);
    const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int hl = 63 - __clzll((long long)(heads & le));
    const unsigned long long gt = heads & ~le;
    const int nextl = gt ? __ffsll((long long)gt) - 1 : 64;
    uint32_t base = 0;
    if (head) {
        // lanes past the last live marker are not live: the run ends there
        int end = nextl;
        const int nlive = (int)st->nm - (k - lane);
        if (end > nlive) end = nlive;
        base = atomicAdd(&tcnt[home], (uint32_t)(end - lane));
    }
    base = __shfl(base, hl < 0 ? 0 : hl, 64);
    if (!live) return;
    tslot[k] = base + (uint32_t)(lane - hl);
    for (int ty = (fy - 1) / ST; ty <= (fy + 2) / ST; ty++)
        for (int tx = (fx - 1) / ST; tx <= (fx + 2) / ST; tx++)
            if (tx < ntx && ty < nty) tact[ty * ntx + tx] = 1;
}

__global__ void k_tile_scatter(int ntx, const double* __restrict__ mx, const double* __restrict__ my,
                               const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tslot, int32_t* tlist,
                               uint32_t* tpos,
                               const DevStatus* st) {
    if (step_stopped(st)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int)st->nm) return;
    const int fx = (int)floor(mx[k]), fy = (int)floor(my[k]);
    const uint32_t e = toff[(fy / ST) * ntx + fx / ST] + tslot[k];
    tlist[e] = k;
    tpos[e] = (uint32_t)fx | ((uint32_t)fy << 16);   // markers lie inside the lattice: 0 <= fx, fy < 65536
}

// One workgroup per tile; tiles no marker stencil reaches return at once.
// The gathered markers are ranked by index and staged SP_CH at a time with
// their eight delta factors; each staged marker sets its bit in a per-site
// mask of the sites its stencil covers in this tile, and every lane then
// walks only the set bits of its site's mask, lowest first, so it adds
// delta*F of exactly the covering markers in rank order.
// F32OUT: mixed-precision lattices take the force as FP32 pairs in Fx's
// allocation (the sums themselves are FP64 either way).
template <bool F32OUT>
__global__ __launch_bounds__(256) void k_tile_spread(int nx, int ny, int ntx, int nty, const uint8_t* __restrict__ tact,
                                                     const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tcnt,
                                                     const int32_t* __restrict__ tlist,
                                                     const uint32_t* __restrict__ tpos, const double* __restrict__ mx,
                                                     const double* __restrict__ my, const double* __restrict__ mfx,
                                                     const double* __restrict__ mfy, double* Fx, double* Fy,
                                                     DevStatus* st) {
    static_assert(SP_CH == 128, "two 64-bit site masks per staged chunk");
    const int t = (int)xcd_contiguous(blockIdx.x, gridDim.x);
    if (!tact[t] || step_stopped(st)) return;
    __shared__ int32_t idx[SP_IDX];   // gathered, then in rank order
    __shared__ double fxa[SP_CH][4], fya[SP_CH][4], sfx[SP_CH], sfy[SP_CH];
    __shared__ int32_t s0x[SP_CH], s0y[SP_CH];
    __shared__ unsigned long long smask[2][ST * ST];
    __shared__ uint32_t hcnt[9], hoff[9], hpre[10];
    __shared__ uint32_t n;
    const int tx = t % ntx, ty = t / ntx;
    const int bx = tx * ST, by = ty * ST;
    // the 3x3 home tiles' lists, loaded at once
    if (threadIdx.x < 9) {
        const int hx = tx + (int)(threadIdx.x % 3) - 1, hy = ty + (int)(threadIdx.x / 3) - 1;
        const bool in = hx >= 0 && hy >= 0 && hx < ntx && hy < nty;
        const int h = hy * ntx + hx;
        hcnt[threadIdx.x] = in ? tcnt[h] : 0u;
        hoff[threadIdx.x] = in ? toff[h] : 0u;
    }
    if (threadIdx.x == 0) n = 0;
    smask[0][threadIdx.x] = 0ull;
    smask[1][threadIdx.x] = 0ull;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0;
        for (int h = 0; h < 9; h++) { hpre[h] = a; a += hcnt[h]; }
        hpre[9] = a;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < hpre[9]; j += blockDim.x) {
        int h = 0;
        while (j >= hpre[h + 1]) h++;
        const uint32_t e = hoff[h] + (j - hpre[h]);
        const int k = tlist[e];
        // keep only markers whose 4x4 stencil overlaps this tile (the binning
        // stored floor(X), floor(Y) with the list entry)
        const uint32_t pp = tpos[e];
        const int fx = (int)(pp & 0xffffu), fy = (int)(pp >> 16);
        if (fx + 2 < bx || fx - 1 > bx + ST - 1 || fy + 2 < by || fy - 1 > by + ST - 1) continue;
        const uint32_t p = atomicAdd(&n, 1u);
        if (p < (uint32_t)SP_IDX) idx[p] = k;
    }
    __syncthreads();
    const uint32_t m = n;
    if (m > (uint32_t)SP_IDX) {
        if (threadIdx.x == 0) atomicOr(&st->bits, LBIB_ERR_CAPACITY);
        return;
    }
    // rank sort in place: ranks into registers, then one scatter (the
    // workgroup is 256 wide, so each lane ranks at most SP_IDX / 256 markers)
    constexpr int RPL = SP_IDX / 256;
    int32_t rv[RPL];
    uint32_t rr[RPL];
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const uint32_t j = threadIdx.x + 256u * q;
        rv[q] = j < m ? idx[j] : 0;
        uint32_t r = 0;
        if (j < m)
            for (uint32_t i = 0; i < m; i++) r += (idx[i] < rv[q]) ? 1u : 0u;
        rr[q] = r;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPL; q++)
        if (threadIdx.x + 256u * q < m) idx[rr[q]] = rv[q];
    __syncthreads();
    const int lx = (int)(threadIdx.x % ST), ly = (int)(threadIdx.x / ST);
    const int px = bx + lx, py = by + ly;
    double ax = 0.0, ay = 0.0;
    for (uint32_t c0 = 0; c0 < m; c0 += SP_CH) {
        const uint32_t mc = min((uint32_t)SP_CH, m - c0);
        for (uint32_t j = threadIdx.x; j < mc; j += blockDim.x) {
            const int k = idx[c0 + j];
            const double X = mx[k], Y = my[k];
            const int x0 = (int)floor(X) - 1, y0 = (int)floor(Y) - 1;
            s0x[j] = x0;
            s0y[j] = y0;
            ib_factors(x0, X, fxa[j]);
            ib_factors(y0, Y, fya[j]);
            sfx[j] = mfx[k];
            sfy[j] = mfy[k];
            const unsigned long long bit = 1ull << (j & 63u);
            for (int qy = 0; qy < 4; qy++)
                for (int qx = 0; qx < 4; qx++) {
                    const int sx = x0 + qx - bx, sy = y0 + qy - by;
                    if ((unsigned)sx < (unsigned)ST && (unsigned)sy < (unsigned)ST)
                        atomicOr(&smask[j >> 6][sy * ST + sx], bit);
                }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            unsigned long long mk = smask[h][threadIdx.x];
            while (mk) {
                const int j = 64 * h + __ffsll((long long)mk) - 1;
                mk &= mk - 1ull;
                const double dd = (fxa[j][px - s0x[j]] * fya[j][py - s0y[j]]) / 16.0;
                ax += dd * sfx[j];
                ay += dd * sfy[j];
            }
            smask[h][threadIdx.x] = 0ull;
        }
        __syncthreads();
    }
    if (px < nx && py < ny) {
        const long site = (long)py * nx + px;
        if (F32OUT) {
            reinterpret_cast<float2*>(Fx)[site] = make_float2((float)ax, (float)ay);
        } else {
            Fx[site] = ax;
            Fy[site] = ay;
        }
    }
}

// GeometryNode::collectVelocity (GeometryNode.cpp:216-226) in stencil-slot
// order, the velocity guard of LagrangianPoint::setVelocity
// (LagrangianPoint.cpp:92-103), GeometryNode::move (GeometryNode.cpp:245-254),
// and detection of the perturbation cases of GeometryNode::perturbatePosition
// (GeometryNode.cpp:153-214) and of the boundary checks of
// GeometryHandler.cpp:613-636.
// UT: storage type of the velocity pairs (FP64, or FP32 for mixed-precision
// lattices); the interpolation and the move are FP64 either way.
template <typename UT>
__global__ void k_interp_move(int nx, int ny, double* mx, double* my, double* mvx,
                              double* mvy, double* mox, double* moy, const UT* __restrict__ u,
                              const int32_t* rtile, const RectDev* rects, int stx, DevStatus* st) {
    if (step_stopped(st)) return;
    // the grid covers the capacity: renumber over the workgroups with markers
    const uint32_t nm = st->nm, nact = (nm + blockDim.x - 1) / blockDim.x;
    if (blockIdx.x >= nact) return;
    const int k = (int)(xcd_contiguous(blockIdx.x, nact) * blockDim.x + threadIdx.x);
    if (k >= (int)nm) return;
    const double X = mx[k], Y = my[k];
    const int fx = (int)floor(X), fy = (int)floor(Y);
    const int base = 4 * (fx - 1) + (fy - 1);
    double fxa[4], fya[4];
    ib_factors(fx - 1, X, fxa);
    ib_factors(fy - 1, Y, fya);
    double vx = 0.0, vy = 0.0;
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const int t = (((s - base) % 16) + 16) % 16;
        const int sx = fx - 1 + (t >> 2), sy = fy - 1 + (t & 3);
        const long site = (long)sy * nx + sx;
        const double dd = (fxa[t >> 2] * fya[t & 3]) / 16.0;
        double ux_s, uy_s;
        if constexpr (sizeof(UT) == 8) {
            const double2 us = *reinterpret_cast<const double2*>(u + 2 * site);
            ux_s = us.x;
            uy_s = us.y;
        } else {
            const float2 us = *reinterpret_cast<const float2*>(u + 2 * site);
            ux_s = (double)us.x;
            uy_s = (double)us.y;
        }
        vx += dd * ux_s;
        vy += dd * uy_s;
    }
    mvx[k] = vx;
    mvy[k] = vy;
    // Constants.hpp:33, CS = 1 / std::sqrt(3) (correctly rounded on the host)
    const double CSv = 0.57735026918962584;
    if (((vx * vx) + (vy * vy)) > CSv) {
        atomicOr(&st->bits, LBIB_ERR_VELOCITY);
        atomicMin(&st->first_bad, k);
    }
    const double Xn = X + vx, Yn = Y + vy;
    mox[k] = X;
    moy[k] = Y;
    mx[k] = Xn;
    my[k] = Yn;
    const bool p1 = fabs((double)(int)Xn - Xn) <= kEPS;
    const bool p2 = fabs((double)(int)Yn - Yn) <= kEPS;
    const bool p3 = fabs(Yn - floor(Yn) - (Xn - floor(Xn))) <= kEPS;
    const bool p4 = fabs(ceil(Yn) - Yn - (Xn - floor(Xn))) <= kEPS;
    if (p1 || p2 || p3 || p4) atomicOr(&st->bits, LBIB_HALT_PERTURB_NODE);
    const bool within = (Xn >= floor(X)) && (Xn < ceil(X)) && (Yn <= ceil(Y)) && (Yn > floor(Y));
    if (!within) {
        // the edges of the node's periodic domain (its rectangle in a batched atlas)
        Frame fr;
        if (!site_frame(rtile, rects, stx, nx, ny, fx, fy, fr)) { fr.x0 = 0; fr.y0 = 0; fr.w = nx; fr.h = ny; }
        if (Xn - 2.0 < fr.x0 || Yn - 2.0 < fr.y0 || Xn + 2.0 > fr.x0 + fr.w || Yn + 2.0 > fr.y0 + fr.h) {
            atomicOr(&st->bits, LBIB_ERR_EDGE);
            atomicMin(&st->first_bad, k);
        }
    }
}

static inline int nblocks(long n, int bs) { return (int)((n + bs - 1) / bs); }

int launch_force_and_spread(lbib_ctx* c) {
    const int bs = 256;
    const int ntiles = c->stx * c->sty;
    // marker counts are device-resident (remeshing changes them): grids cover
    // the capacity and the kernels read st->nm
    const int begin = c->begin_pending ? 1 : 0;
    c->begin_pending = false;
    k_marker_force<<<std::max(nblocks(c->mcap, bs), 1), bs, 0, c->stream>>>(
        c->mx, c->my, c->fs_off, c->fs_type, c->fs_partner, c->fs_prm, c->mfx, c->mfy, ntiles, c->tcnt, c->tact, begin,
        c->capturing ? ITER_NEXT : c->begin_it, c->st);
    if (c->nm == 0) {
        LBIB_HIP(hipGetLastError());
        return 0;
    }
    k_tile_count<<<nblocks(c->mcap, bs), bs, 0, c->stream>>>(c->stx, c->sty, c->mx, c->my, c->tcnt, c->tslot,
                                                             c->tact, c->st);
    if (scan_tiles(c, c->tact, &c->st->ntouched)) return -1;
    k_tile_scatter<<<nblocks(c->mcap, bs), bs, 0, c->stream>>>(c->stx, c->mx, c->my, c->toff, c->tslot, c->tlist, c->tpos,
                                                               c->st);
    if (c->prec == LBIB_FP64)
        k_tile_spread<false><<<ntiles, 256, 0, c->stream>>>(c->nx, c->ny, c->stx, c->sty, c->tact, c->toff, c->tcnt,
                                                            c->tlist, c->tpos, c->mx, c->my, c->mfx, c->mfy, c->Fx, c->Fy, c->st);
    else
        k_tile_spread<true><<<ntiles, 256, 0, c->stream>>>(c->nx, c->ny, c->stx, c->sty, c->tact, c->toff, c->tcnt,
                                                           c->tlist, c->tpos, c->mx, c->my, c->mfx, c->mfy, c->Fx, c->Fy, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

__global__ void k_tile_clear(uint32_t* tcnt, uint32_t n, const DevStatus* st) {
    if (step_stopped(st)) return;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) tcnt[t] = 0u;
}

int launch_tile_bins(lbib_ctx* c) {
    const int bs = 256;
    const uint32_t nt = (uint32_t)c->stx * (uint32_t)c->sty;
    k_tile_clear<<<std::min(nblocks((long)nt, bs), 1024), bs, 0, c->stream>>>(c->tcnt, nt, c->st);
    k_tile_count<<<nblocks(c->mcap, bs), bs, 0, c->stream>>>(c->stx, c->sty, c->mx, c->my, c->tcnt, c->tslot,
                                                             c->tact, c->st);
    if (scan_exclusive(c, c->tcnt, c->toff, nullptr, nt, nt, nullptr)) return -1;
    k_tile_scatter<<<nblocks(c->mcap, bs), bs, 0, c->stream>>>(c->stx, c->mx, c->my, c->toff, c->tslot, c->tlist, c->tpos,
                                                               c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_clear_force(lbib_ctx*) { return 0; }

int launch_interp_move(lbib_ctx* c) {
    if (c->nm == 0) return 0;
    const int32_t* rt = c->nrect ? c->rtile : nullptr;
    if (c->prec == LBIB_FP64)
        k_interp_move<double><<<nblocks(c->mcap, 256), 256, 0, c->stream>>>(
            c->nx, c->ny, c->mx, c->my, c->mvx, c->mvy, c->mox, c->moy, c->u, rt, c->rects, c->stx, c->st);
    else
        k_interp_move<float><<<nblocks(c->mcap, 256), 256, 0, c->stream>>>(
            c->nx, c->ny, c->mx, c->my, c->mvx, c->mvy, c->mox, c->moy, reinterpret_cast<const float*>(c->u), rt,
            c->rects, c->stx, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

}  // namespace lbib
