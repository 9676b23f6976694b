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
#include "lbib_device_math.h"

namespace lbib {

// ForceSolver::calculateForce (ForceSolver.cpp:42-85) with the force structs of
// ForceStructs.hpp:49-588; per marker, structs in map/push order.
__global__ void k_marker_force(int nm, const double* __restrict__ mx, const double* __restrict__ my,
                               const int32_t* __restrict__ off, const int32_t* __restrict__ type,
                               const int32_t* __restrict__ partner, const double* __restrict__ prm,
                               double* __restrict__ mfx, double* __restrict__ mfy) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
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

__device__ __forceinline__ void stencil_site(const double X, const double Y, const int a, const int b,
                                             int& sx, int& sy) {
    sx = (int)floor(X) - 1 + a;
    sy = (int)floor(Y) - 1 + b;
}

// pass 1: count contributions per site; first toucher appends the site and
// raises its tile flag.
__global__ void k_spread_count(int nm, int nx, int ntx, const double* __restrict__ mx,
                               const double* __restrict__ my, uint32_t* scnt, uint32_t* srank,
                               uint32_t* touched, uint8_t* ftile, DevStatus* st) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nm * 16) return;
    const int k = (int)(t >> 4), s = (int)(t & 15);
    int sx, sy;
    stencil_site(mx[k], my[k], s >> 2, s & 3, sx, sy);
    const long site = (long)sy * nx + sx;
    const uint32_t r = atomicAdd(&scnt[site], 1u);
    srank[t] = r;
    if (r == 0) {
        const uint32_t p = atomicAdd(&st->ntouched, 1u);
        touched[p] = (uint32_t)site;
        ftile[(long)sy * ntx + sx / TILE_W] = 1;
    }
}

// pass 2: allocate a pool range per touched site.
__global__ void k_spread_alloc(const uint32_t* __restrict__ touched, const uint32_t* __restrict__ scnt,
                               uint32_t* sbase, DevStatus* st) {
    const uint32_t nt = st->ntouched;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        const uint32_t site = touched[t];
        sbase[site] = atomicAdd(&st->pool_used, scnt[site]);
    }
}

// pass 3: scatter (marker, delta*F) records.  Same arithmetic as
// GeometryNode::distributeForce (GeometryNode.cpp:228-235).
__global__ void k_spread_scatter(int nm, int nx, const double* __restrict__ mx,
                                 const double* __restrict__ my, const double* __restrict__ mfx,
                                 const double* __restrict__ mfy, const uint32_t* __restrict__ srank,
                                 const uint32_t* __restrict__ sbase, SpreadEntry* pool) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nm * 16) return;
    const int k = (int)(t >> 4), s = (int)(t & 15);
    const double X = mx[k], Y = my[k];
    int sx, sy;
    stencil_site(X, Y, s >> 2, s & 3, sx, sy);
    const long site = (long)sy * nx + sx;
    const double dd = ib_delta((double)sx - X, (double)sy - Y);
    SpreadEntry e;
    e.marker = k;
    e.pad = 0;
    e.cx = dd * mfx[k];
    e.cy = dd * mfy[k];
    pool[sbase[site] + srank[t]] = e;
}

// pass 4: per touched site, sum in marker order (FluidSolver::addForce,
// FluidSolver.cpp:196-198, starting from resetForce's 0.0).
__global__ void k_spread_reduce(const uint32_t* __restrict__ touched, uint32_t* scnt,
                                const uint32_t* __restrict__ sbase, SpreadEntry* pool, double* Fx,
                                double* Fy, DevStatus* st) {
    const uint32_t nt = st->ntouched;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        const uint32_t site = touched[t];
        const uint32_t cnt = scnt[site];
        SpreadEntry* e = pool + sbase[site];
        // insertion sort by marker index (counts are small: markers within 2 lu)
        for (uint32_t a = 1; a < cnt; a++) {
            SpreadEntry v = e[a];
            int b = (int)a - 1;
            while (b >= 0 && e[b].marker > v.marker) { e[b + 1] = e[b]; b--; }
            e[b + 1] = v;
        }
        double sx = 0.0, sy = 0.0;
        for (uint32_t a = 0; a < cnt; a++) { sx += e[a].cx; sy += e[a].cy; }
        Fx[site] = sx;
        Fy[site] = sy;
        scnt[site] = 0;
    }
}

// after the collide: zero the force of touched sites and their tile flags.
__global__ void k_clear_force(const uint32_t* __restrict__ touched, double* Fx, double* Fy, int nx,
                              int ntx, uint8_t* ftile, DevStatus* st) {
    const uint32_t nt = st->ntouched;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        const uint32_t site = touched[t];
        Fx[site] = 0.0;
        Fy[site] = 0.0;
        ftile[(site / nx) * (long)ntx + (site % nx) / TILE_W] = 0;
    }
}

__global__ void k_reset_spread_counters(DevStatus* st) {
    st->ntouched = 0;
    st->pool_used = 0;
}

// GeometryNode::collectVelocity (GeometryNode.cpp:216-226) in stencil-slot
// order, the velocity guard of LagrangianPoint::setVelocity
// (LagrangianPoint.cpp:92-103), GeometryNode::move (GeometryNode.cpp:245-254),
// and detection of the perturbation cases of GeometryNode::perturbatePosition
// (GeometryNode.cpp:153-214) and of the boundary checks of
// GeometryHandler.cpp:613-636.
__global__ void k_interp_move(int nm, int nx, int ny, double* mx, double* my, double* mvx,
                              double* mvy, double* mox, double* moy, const double* __restrict__ ux,
                              const double* __restrict__ uy, DevStatus* st) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const double X = mx[k], Y = my[k];
    const int fx = (int)floor(X), fy = (int)floor(Y);
    const int base = 4 * (fx - 1) + (fy - 1);
    double vx = 0.0, vy = 0.0;
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const int t = (((s - base) % 16) + 16) % 16;
        const int sx = fx - 1 + (t >> 2), sy = fy - 1 + (t & 3);
        const long site = (long)sy * nx + sx;
        const double dd = ib_delta((double)sx - X, (double)sy - Y);
        vx += dd * ux[site];
        vy += dd * uy[site];
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
        if (Xn - 2.0 < 0 || Yn - 2.0 < 0 || Xn + 2.0 > nx || Yn + 2.0 > ny) {
            atomicOr(&st->bits, LBIB_ERR_EDGE);
            atomicMin(&st->first_bad, k);
        }
    }
}

static inline int nblocks(long n, int bs) { return (int)((n + bs - 1) / bs); }

int launch_force_and_spread(lbib_ctx* c) {
    if (c->nm == 0) return 0;
    const int bs = 256;
    k_marker_force<<<nblocks(c->nm, bs), bs, 0, c->stream>>>(c->nm, c->mx, c->my, c->fs_off, c->fs_type,
                                                             c->fs_partner, c->fs_prm, c->mfx, c->mfy);
    k_reset_spread_counters<<<1, 1, 0, c->stream>>>(c->st);
    const long nc = (long)c->nm * 16;
    k_spread_count<<<nblocks(nc, bs), bs, 0, c->stream>>>(c->nm, c->nx, c->ntx, c->mx, c->my, c->scnt,
                                                          c->srank, c->touched, c->ftile, c->st);
    const int gs = 1024;
    k_spread_alloc<<<gs, bs, 0, c->stream>>>(c->touched, c->scnt, c->sbase, c->st);
    k_spread_scatter<<<nblocks(nc, bs), bs, 0, c->stream>>>(c->nm, c->nx, c->mx, c->my, c->mfx, c->mfy,
                                                            c->srank, c->sbase, c->pool);
    k_spread_reduce<<<gs, bs, 0, c->stream>>>(c->touched, c->scnt, c->sbase, c->pool, c->Fx, c->Fy, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_clear_force(lbib_ctx* c) {
    if (c->nm == 0) return 0;
    k_clear_force<<<1024, 256, 0, c->stream>>>(c->touched, c->Fx, c->Fy, c->nx, c->ntx, c->ftile, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_interp_move(lbib_ctx* c) {
    if (c->nm == 0) return 0;
    k_interp_move<<<nblocks(c->nm, 256), 256, 0, c->stream>>>(c->nm, c->nx, c->ny, c->mx, c->my, c->mvx,
                                                             c->mvy, c->mox, c->moy, c->ux, c->uy, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

}  // namespace lbib
