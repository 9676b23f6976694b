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
// Ordered spread (FP64 lattices).  The reference adds delta*F into each site
// in marker-id order (SimulationRunner.cpp:449-452).  To reproduce those FP64
// sums exactly, markers are binned by the 16x16 tile of their position (count,
// scan, scatter), and one workgroup per tile gathers the markers of the 3x3
// tiles around it, ranks them by index and adds each site's contributions in
// that order.  Atomics only count; no floating-point atomics.  Mixed-precision
// lattices take an order-free fixed-point sum instead (see k_spread_fixed).
#include <algorithm>
#include <climits>
#include "lbib_device_math.h"

namespace lbib {

// ForceSolver::calculateForce (ForceSolver.cpp:42-85) with the force structs of
// ForceStructs.hpp:49-588; per marker, structs in map/push order.
__device__ __forceinline__ void marker_force(const int32_t* __restrict__ off, const int32_t* __restrict__ type,
                                             const int32_t* __restrict__ partner, const double* __restrict__ prm,
                                             const double* __restrict__ mx, const double* __restrict__ my, int k,
                                             double x1, double y1, double& fx, double& fy) {
    fx = 0.0;
    fy = 0.0;
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
}

// ---- marker bins over 16x16 site tiles -----------------------------------
// A tile's sites can be reached only by markers whose stencil overlaps it.
// Two kinds of bins: ordered bins list each marker once, in the tile of
// floor(X), floor(Y) (the FP64 spread gathers the 3x3 tiles around a tile and
// the junction search the tiles within its radius); push bins list each
// marker in every tile its stencil reaches, with what the order-free spread
// needs.  Each tile holds BINCAP entries; a slot is taken with one atomic per
// run of lanes binning into the same tile (consecutive markers of a membrane
// mostly share one), and entries past the capacity go to the overflow list.
constexpr int ST = 16;            // spread tile edge
constexpr int SP_IDX = 1024;      // markers gathered per tile
constexpr int SP_CH = 64;         // marker data staged per pass (one 64-bit site mask)
constexpr int SP_MW = SP_CH / 64;  // mask words per site
constexpr int SPW = 128;           // workgroup width of the ordered spread
constexpr int SPS = ST * ST / SPW; // sites of the tile per lane

// Clears the bins: tile counts, active flags and the overflow count.  Run
// by itself before a binning pass that does not follow an interpolation,
// and by every interpolation (k_interp_move) for the next step's binning,
// also in a stopped step (a rerun then bins into empty tiles).
__device__ __forceinline__ void clear_bins(uint32_t* tcnt, uint8_t* tact, int nt, DevStatus* st) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        tcnt[t] = 0u;
        tact[t] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->novf = 0u;
}
__global__ void k_bin_clear(uint32_t* tcnt, uint8_t* tact, int nt, DevStatus* st) { clear_bins(tcnt, tact, nt, st); }

// The slot of this lane's entry in tile `key`: the head lane of each run of
// equal keys in the wavefront takes the run's slots with one atomic.  Lanes
// that bin nothing pass a key of their own (-1 - lane), which ends every run.
// Called by the whole wavefront.
__device__ __forceinline__ uint32_t run_slot(int key, bool part, uint32_t* tcnt) {
    const int lane = threadIdx.x & 63;
    const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int kprev = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || kprev != key;
    const unsigned long long heads = __ballot(head);
    const int hl = 63 - __clzll((long long)(heads & le));
    const unsigned long long gt = heads & ~le;
    const int nextl = gt ? __ffsll((long long)gt) - 1 : 64;
    uint32_t base = 0;
    if (head && part) base = atomicAdd(&tcnt[key], (uint32_t)(nextl - lane));
    base = __shfl(base, hl, 64);
    return base + (uint32_t)(lane - hl);
}

__device__ __forceinline__ void bin_overflow(BinOvf* ovf, DevStatus* st, const BinOvf& e) {
    const uint32_t o = atomicAdd(&st->novf, 1u);
    if (o < OVF_CAP) ovf[o] = e;
    else atomicOr(&st->bits, LBIB_ERR_CAPACITY);
}

// The tiles a marker's 4x4 stencil reaches: quadrant q takes the tile of the
// stencil's corner (q & 1: right column, q & 2: top row) when it differs from
// the tiles of the lower quadrants.
struct StencilTiles {
    int tx0, tx1, ty0, ty1;
    __device__ StencilTiles(int x0, int y0) : tx0(x0 / ST), tx1((x0 + 3) / ST), ty0(y0 / ST), ty1((y0 + 3) / ST) {}
    __device__ bool has(int q) const { return ((q & 1) == 0 || tx1 != tx0) && ((q & 2) == 0 || ty1 != ty0); }
    __device__ int tile(int q, int ntx) const { return ((q & 2) ? ty1 : ty0) * ntx + ((q & 1) ? tx1 : tx0); }
};

struct BinArgs {
    const double *mx, *my;
    const int32_t *off, *type, *partner;   // force CSR; off == nullptr: bin only
    const double* prm;
    double *mfx, *mfy;
    int ntx, nty;
    uint32_t* tcnt;
    uint8_t* tact;
    int32_t* tlist;                        // ordered bins
    uint32_t* tpos;
    uint32_t* pb_pos;                      // push bins
    float2 *pb_r, *pb_f;
    BinOvf* ovf;
    uint32_t* shard;
};

// The marker forces and the binning of the lattice stage in one pass over the
// markers (PUSH: push bins for the order-free spread, else ordered bins).
// When it runs on the main stream it also opens a pending iteration
// (lbib_step_begin).  Bins must be clear (k_bin_clear).
template <int PUSH>
__global__ __launch_bounds__(256) void k_force_bin(const BinArgs a, int begin, uint32_t begin_it, DevStatus* st) {
    if (begin && blockIdx.x == 0 && threadIdx.x == 0) step_begin_dev(st, begin_it);
    if (step_stopped(st)) return;
    // the touched-tile count of the spread that follows
    if (blockIdx.x == 0 && threadIdx.x < NSHARD) a.shard[(SH_TOUCHED * NSHARD + threadIdx.x) * SHARD_STRIDE] = 0u;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = k < (int)st->nm;
    const int lane = threadIdx.x & 63;
    double X = 2.0, Y = 2.0, fx = 0.0, fy = 0.0;
    if (live) {
        X = a.mx[k];
        Y = a.my[k];
        if (a.off) {
            marker_force(a.off, a.type, a.partner, a.prm, a.mx, a.my, k, X, Y, fx, fy);
            a.mfx[k] = fx;
            a.mfy[k] = fy;
        }
    }
    // markers lie inside the lattice (the uploads and the move check the
    // edges): 0 <= floor(X), floor(Y) < 65536
    const int ix = (int)floor(X), iy = (int)floor(Y);
    if (!PUSH) {
        const int key = live ? (iy / ST) * a.ntx + ix / ST : -1 - lane;
        const uint32_t slot = run_slot(key, live, a.tcnt);
        if (!live) return;
        const uint32_t pos = (uint32_t)ix | ((uint32_t)iy << 16);
        if (slot < (uint32_t)BINCAP) {
            const long e = (long)key * BINCAP + slot;
            a.tlist[e] = k;
            a.tpos[e] = pos;
        } else {
            bin_overflow(a.ovf, st, BinOvf{key, pos, k, 0.0f, 0.0f, 0.0f, 0.0f});
        }
        for (int ty = (iy - 1) / ST; ty <= (iy + 2) / ST; ty++)
            for (int tx = (ix - 1) / ST; tx <= (ix + 2) / ST; tx++)
                if (tx < a.ntx && ty < a.nty) a.tact[ty * a.ntx + tx] = 1;
    } else {
        const int x0 = ix - 1, y0 = iy - 1;
        const StencilTiles T(x0, y0);
        const uint32_t pos = (uint32_t)x0 | ((uint32_t)y0 << 16);
        const float2 r = make_float2((float)(X - (double)x0), (float)(Y - (double)y0));
        const float2 f = make_float2((float)fx, (float)fy);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const bool part = live && T.has(q);
            const int key = part ? T.tile(q, a.ntx) : -1 - lane;
            const uint32_t slot = run_slot(key, part, a.tcnt);
            if (!part) continue;
            a.tact[key] = 1;
            if (slot < (uint32_t)BINCAP) {
                const long e = (long)key * BINCAP + slot;
                a.pb_pos[e] = pos;
                a.pb_r[e] = r;
                a.pb_f[e] = f;
            } else {
                bin_overflow(a.ovf, st, BinOvf{key, pos, k, r.x, r.y, f.x, f.y});
            }
        }
    }
}

// Shared staging of one tile's spread.
struct SpreadLds {
    int32_t idx[SP_IDX];   // gathered, then in rank order
    double fxa[SP_CH][4], fya[SP_CH][4], sfx[SP_CH], sfy[SP_CH];
    int32_t s0x[SP_CH], s0y[SP_CH];
    unsigned long long smask[SP_MW][ST * ST];
    uint32_t hcnt[9], hpre[10];
    int32_t htile[9];
    uint32_t n, novf;
};

// One tile of the ordered spread.  The gathered markers are ranked by index
// and staged SP_CH at a time with their eight delta factors; each staged
// marker sets its bit in a per-site mask of the sites its stencil covers in
// this tile, and every lane then walks only the set bits of its sites' masks
// (SPS sites per lane), lowest first, so it adds delta*F of exactly the
// covering markers in rank order.  Called by the whole workgroup of SPW lanes
// with a workgroup-uniform tile; returns false when the tile's marker count
// exceeds SP_IDX.  A 128-lane workgroup with 10 KB of LDS lets 16 tiles
// reside per CU.
__device__ __forceinline__ bool spread_tile(SpreadLds& L, int t, int nx, int ny, int ntx, int nty,
                                            const uint32_t* __restrict__ tcnt, const int32_t* __restrict__ tlist,
                                            const uint32_t* __restrict__ tpos, const BinOvf* __restrict__ ovf,
                                            const double* __restrict__ mx, const double* __restrict__ my,
                                            const double* __restrict__ mfx, const double* __restrict__ mfy,
                                            double* Fx, double* Fy, const DevStatus* st) {
    const int tx = t % ntx, ty = t / ntx;
    const int bx = tx * ST, by = ty * ST;
    // the 3x3 home tiles' bins, counts loaded at once
    if (threadIdx.x < 9) {
        const int hx = tx + (int)(threadIdx.x % 3) - 1, hy = ty + (int)(threadIdx.x / 3) - 1;
        const bool in = hx >= 0 && hy >= 0 && hx < ntx && hy < nty;
        const int h = hy * ntx + hx;
        L.hcnt[threadIdx.x] = in ? tcnt[h] : 0u;
        L.htile[threadIdx.x] = in ? h : -1;
    }
    if (threadIdx.x == 0) {
        L.n = 0;
        L.novf = min(st->novf, OVF_CAP);
    }
#pragma unroll
    for (int u = 0; u < SPS; u++)
#pragma unroll
        for (int h = 0; h < SP_MW; h++) L.smask[h][threadIdx.x + SPW * u] = 0ull;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0;
        bool over = false;
        for (int h = 0; h < 9; h++) {
            L.hpre[h] = a;
            a += min(L.hcnt[h], (uint32_t)BINCAP);
            over |= L.hcnt[h] > (uint32_t)BINCAP;
        }
        L.hpre[9] = a;
        if (!over) L.novf = 0;   // none of these tiles overflowed
    }
    __syncthreads();
    // keep only markers whose 4x4 stencil overlaps this tile (the binning
    // stored floor(X), floor(Y) with the entry)
    auto gather = [&](int k, uint32_t pp) {
        const int fx = (int)(pp & 0xffffu), fy = (int)(pp >> 16);
        if (fx + 2 < bx || fx - 1 > bx + ST - 1 || fy + 2 < by || fy - 1 > by + ST - 1) return;
        const uint32_t p = atomicAdd(&L.n, 1u);
        if (p < (uint32_t)SP_IDX) L.idx[p] = k;
    };
    for (uint32_t j = threadIdx.x; j < L.hpre[9]; j += blockDim.x) {
        int h = 0;
        while (j >= L.hpre[h + 1]) h++;
        const long e = (long)L.htile[h] * BINCAP + (j - L.hpre[h]);
        gather(tlist[e], tpos[e]);
    }
    for (uint32_t j = threadIdx.x; j < L.novf; j += blockDim.x) {
        const BinOvf o = ovf[j];
        bool mine = false;
        for (int h = 0; h < 9; h++) mine |= o.tile == L.htile[h];
        if (mine) gather(o.k, o.pos);
    }
    __syncthreads();
    const uint32_t m = L.n;
    if (m > (uint32_t)SP_IDX) return false;
    // rank sort in place: ranks into registers, then one scatter (each lane
    // ranks at most SP_IDX / SPW markers)
    constexpr int RPL = SP_IDX / SPW;
    int32_t rv[RPL];
    uint32_t rr[RPL];
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const uint32_t j = threadIdx.x + (uint32_t)SPW * q;
        rv[q] = j < m ? L.idx[j] : 0;
        uint32_t r = 0;
        if (j < m)
            for (uint32_t i = 0; i < m; i++) r += (L.idx[i] < rv[q]) ? 1u : 0u;
        rr[q] = r;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPL; q++)
        if (threadIdx.x + (uint32_t)SPW * q < m) L.idx[rr[q]] = rv[q];
    __syncthreads();
    // each lane sums SPS sites of the tile (independent chains)
    double ax[SPS], ay[SPS];
#pragma unroll
    for (int u = 0; u < SPS; u++) {
        ax[u] = 0.0;
        ay[u] = 0.0;
    }
    for (uint32_t c0 = 0; c0 < m; c0 += SP_CH) {
        const uint32_t mc = min((uint32_t)SP_CH, m - c0);
        for (uint32_t j = threadIdx.x; j < mc; j += blockDim.x) {
            const int k = L.idx[c0 + j];
            const double X = mx[k], Y = my[k];
            const int x0 = (int)floor(X) - 1, y0 = (int)floor(Y) - 1;
            L.s0x[j] = x0;
            L.s0y[j] = y0;
            ib_factors(x0, X, L.fxa[j]);
            ib_factors(y0, Y, L.fya[j]);
            L.sfx[j] = mfx[k];
            L.sfy[j] = mfy[k];
            const unsigned long long bit = 1ull << (j & 63u);
            for (int qy = 0; qy < 4; qy++)
                for (int qx = 0; qx < 4; qx++) {
                    const int sx = x0 + qx - bx, sy = y0 + qy - by;
                    if ((unsigned)sx < (unsigned)ST && (unsigned)sy < (unsigned)ST)
                        atomicOr(&L.smask[j >> 6][sy * ST + sx], bit);
                }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < SPS; u++) {
            const int sl = (int)threadIdx.x + SPW * u;
            const int px = bx + sl % ST, py = by + sl / ST;
#pragma unroll
            for (int h = 0; h < SP_MW; h++) {
                unsigned long long mk = L.smask[h][sl];
                while (mk) {
                    const int j = 64 * h + __ffsll((long long)mk) - 1;
                    mk &= mk - 1ull;
                    const double dd = (L.fxa[j][px - L.s0x[j]] * L.fya[j][py - L.s0y[j]]) / 16.0;
                    ax[u] += dd * L.sfx[j];
                    ay[u] += dd * L.sfy[j];
                }
                L.smask[h][sl] = 0ull;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < SPS; u++) {
        const int sl = (int)threadIdx.x + SPW * u;
        const int px = bx + sl % ST, py = by + sl / ST;
        if (px < nx && py < ny) {
            const long site = (long)py * nx + px;
            Fx[site] = ax[u];
            Fy[site] = ay[u];
        }
    }
    return true;
}

// One workgroup per tile; tiles no marker stencil reaches return at once.
// MAP 0: consecutive tiles on one XCD (xcd_contiguous), for one lattice whose
// active tiles spread evenly; MAP 1: tile rows dealt to the XCDs in turn
// (row_interleaved_tile), for a batched atlas whose occupied rectangles would
// otherwise load some XCDs far more than others.
template <int MAP>
__global__ __launch_bounds__(SPW) void k_tile_spread(int nx, int ny, int ntx, int nty, const uint8_t* __restrict__ tact,
                                                     const uint32_t* __restrict__ tcnt, const int32_t* __restrict__ tlist,
                                                     const uint32_t* __restrict__ tpos, const BinOvf* __restrict__ ovf,
                                                     const double* __restrict__ mx, const double* __restrict__ my,
                                                     const double* __restrict__ mfx, const double* __restrict__ mfy,
                                                     double* Fx, double* Fy, uint32_t* shard, int begin,
                                                     uint32_t begin_it, DevStatus* st) {
    static_assert(SP_CH % 64 == 0, "whole 64-bit site masks per staged chunk");
    if (begin && blockIdx.x == 0 && threadIdx.x == 0) step_begin_dev(st, begin_it);
    const int t = MAP == 0 ? (int)xcd_contiguous(blockIdx.x, gridDim.x) : row_interleaved_tile(blockIdx.x, ntx, nty);
    if (t < 0 || !tact[t] || step_stopped(st)) return;
    if (threadIdx.x == 0) atomicAdd(&shard[(SH_TOUCHED * NSHARD + (blockIdx.x % NSHARD)) * SHARD_STRIDE], 1u);
    __shared__ SpreadLds L;
    if (!spread_tile(L, t, nx, ny, ntx, nty, tcnt, tlist, tpos, ovf, mx, my, mfx, mfy, Fx, Fy, st) &&
        threadIdx.x == 0)
        atomicOr(&st->bits, LBIB_ERR_CAPACITY);
}

// ---- mixed-precision spread ------------------------------------------------
// Lattices stored in FP32 or 16-bit formats do not reproduce the reference's
// FP64 sums anyway, so their force field is summed without an order: each
// tile lists every marker whose stencil reaches it (with its position and
// force), and one wavefront per tile adds the FP32 products delta*F as int64
// multiples of 2^-36 with LDS atomics.  Integer addition is associative, so
// the result does not depend on the order of the atomics (runs are
// reproducible); the quantum 2^-36 is far below FP32 resolution of the
// forces involved.  A wavefront per tile (5 KB of LDS) lets 32 tiles reside
// on a CU at once: the pass over the lattice's tiles is bound by how many
// are in flight, not by their arithmetic.
constexpr float kFixScale = 68719476736.0f;    // 2^36
constexpr double kFixInv = 1.0 / 68719476736.0;

// fastCos (GeometryNode.cpp:28-45) in FP32
__device__ __forceinline__ float fast_cos_f32(float x) {
    float t = x + 1.57079632679f;
    if (t > 3.14159265359f) t -= 6.28318530718f;
    float c = t < 0.0f ? 1.27323954f * t + 0.405284735f * t * t : 1.27323954f * t - 0.405284735f * t * t;
    c = c < 0.0f ? 0.225f * (c * -c - c) + c : 0.225f * (c * c - c) + c;
    return c;
}

template <int MAP>
__global__ __launch_bounds__(64) void k_spread_fixed(int nx, int ny, int ntx, int nty, const uint32_t* __restrict__ tcnt,
                                                      const uint32_t* __restrict__ pb_pos, const float2* __restrict__ pb_r,
                                                      const float2* __restrict__ pb_f, const BinOvf* __restrict__ ovf,
                                                      float2* F, uint32_t* shard, int begin, uint32_t begin_it,
                                                      DevStatus* st) {
    if (begin && blockIdx.x == 0 && threadIdx.x == 0) step_begin_dev(st, begin_it);
    const int t = MAP == 0 ? (int)xcd_contiguous(blockIdx.x, gridDim.x) : row_interleaved_tile(blockIdx.x, ntx, nty);
    if (t < 0 || step_stopped(st)) return;
    const uint32_t n = tcnt[t];
    if (n == 0) return;
    if (threadIdx.x == 0) atomicAdd(&shard[(SH_TOUCHED * NSHARD + (blockIdx.x % NSHARD)) * SHARD_STRIDE], 1u);
    __shared__ unsigned long long acc[2][ST * ST];
    __shared__ uint32_t bad[ST * ST];   // a non-finite or out-of-range product reached the site
    for (int s = threadIdx.x; s < ST * ST; s += blockDim.x) {
        acc[0][s] = 0ull;
        acc[1][s] = 0ull;
        bad[s] = 0u;
    }
    __syncthreads();
    const int tx = t % ntx, ty = t / ntx;
    const int bx = tx * ST, by = ty * ST;
    // one entry: stencil corner (x0, y0), offsets (rx, ry) = (X - x0, Y - y0)
    // and the force, all as the binning rounded them
    auto add = [&](uint32_t pos, float2 r, float2 f) {
        const int x0 = (int)(pos & 0xffffu), y0 = (int)(pos >> 16);
        float fa[4], fb[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            fa[q] = 1.0f + fast_cos_f32(1.57079632679f * ((float)q - r.x));
            fb[q] = 1.0f + fast_cos_f32(1.57079632679f * ((float)q - r.y));
        }
        // delta = fa * fb / 16: the 1/16 goes into the scale
        const float gx = f.x * (kFixScale / 16.0f), gy = f.y * (kFixScale / 16.0f);
#pragma unroll
        for (int qy = 0; qy < 4; qy++) {
            const int sy = y0 + qy - by;
            if ((unsigned)sy >= (unsigned)ST) continue;
#pragma unroll
            for (int qx = 0; qx < 4; qx++) {
                const int sx = x0 + qx - bx;
                if ((unsigned)sx >= (unsigned)ST) continue;
                const float w = fa[qx] * fb[qy];
                const float vx = w * gx, vy = w * gy;
                const int site = sy * ST + sx;
                if (!(fabsf(vx) < 4.0e18f && fabsf(vy) < 4.0e18f)) {
                    bad[site] = 1u;
                    continue;
                }
                atomicAdd(&acc[0][site], (unsigned long long)__float2ll_rn(vx));
                atomicAdd(&acc[1][site], (unsigned long long)__float2ll_rn(vy));
            }
        }
    };
    const long o = (long)t * BINCAP;
    const uint32_t nb = min(n, (uint32_t)BINCAP);
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) add(pb_pos[o + i], pb_r[o + i], pb_f[o + i]);
    if (n > (uint32_t)BINCAP) {
        const uint32_t no = min(st->novf, OVF_CAP);
        for (uint32_t j = threadIdx.x; j < no; j += blockDim.x) {
            const BinOvf e = ovf[j];
            if (e.tile == t) add(e.pos, make_float2(e.rx, e.ry), make_float2(e.fx, e.fy));
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < ST * ST; s += blockDim.x) {
        const int px = bx + s % ST, py = by + s / ST;
        if (px >= nx || py >= ny) continue;
        float2 v;
        if (bad[s]) {
            v = make_float2(__int_as_float(0x7fc00000), __int_as_float(0x7fc00000));
        } else {
            v = make_float2((float)((double)(long long)acc[0][s] * kFixInv), (float)((double)(long long)acc[1][s] * kFixInv));
        }
        F[(long)py * nx + px] = v;
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
                              const int32_t* rtile, const RectDev* rects, int stx, uint32_t* tcnt, uint8_t* tact,
                              int nt, DevStatus* st) {
    // the lattice kernel has read the active-tile flags: the bins are free
    clear_bins(tcnt, tact, nt, st);
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

static BinArgs bin_args(lbib_ctx* c, bool forces) {
    BinArgs a;
    a.mx = c->mx; a.my = c->my;
    a.off = forces ? c->fs_off : nullptr;
    a.type = c->fs_type; a.partner = c->fs_partner; a.prm = c->fs_prm;
    a.mfx = c->mfx; a.mfy = c->mfy;
    a.ntx = c->stx; a.nty = c->sty;
    a.tcnt = c->tcnt; a.tact = c->tact;
    a.tlist = c->tlist; a.tpos = c->tpos;
    a.pb_pos = c->pb_pos; a.pb_r = c->pb_r; a.pb_f = c->pb_f;
    a.ovf = c->ovf;
    a.shard = c->shard;
    return a;
}

// empties the bins unless the last kernel to touch them cleared them
static void launch_bin_clear(lbib_ctx* c) {
    if (c->bins_clear) return;
    const int nt = c->stx * c->sty;
    k_bin_clear<<<std::max(std::min(nblocks(nt, 256), 1024), 1), 256, 0, c->stream>>>(c->tcnt, c->tact, nt, c->st);
}

// The marker forces and the tile binning read the topology, the marker
// positions and the force structs only.  When nothing enqueued since the last
// geometry stage changed those (forces_beside), they follow the topology stage
// on the second stream, beside the main stream's cut links, re-initialisation
// and scanline; the main stream waits for them before the spread, whose first
// workgroup then opens the iteration (after every stop of the last one is
// raised).  Otherwise all of it runs on the main stream, the force kernel
// opening the iteration.
int launch_force_and_spread(lbib_ctx* c) {
    const int bs = 256;
    const int ntiles = c->stx * c->sty;
    const int begin = c->begin_pending ? 1 : 0;
    const uint32_t begin_it = c->capturing ? ITER_NEXT : c->begin_it;
    c->begin_pending = false;
    const bool beside = c->forces_beside && c->nm > 0;
    c->forces_beside = false;
    if (!beside && join_topology(c)) return -1;
    const hipStream_t main = c->stream;
    if (beside) c->stream = c->stream2;
    const bool mixed = c->prec != LBIB_FP64;
    // marker counts are device-resident (remeshing changes them): grids cover
    // the capacity and the kernels read st->nm
    launch_bin_clear(c);
    c->bins_clear = false;
    const BinArgs a = bin_args(c, true);
    const int fgrid = std::max(nblocks(c->mcap, bs), 1);
    if (mixed) k_force_bin<1><<<fgrid, bs, 0, c->stream>>>(a, beside ? 0 : begin, begin_it, c->st);
    else k_force_bin<0><<<fgrid, bs, 0, c->stream>>>(a, beside ? 0 : begin, begin_it, c->st);
    c->stream = main;
    if (c->nm == 0) {
        LBIB_HIP(hipGetLastError());
        return 0;
    }
    // a batched atlas deals tile rows to the XCDs in turn (its occupied
    // rectangles are unevenly spread); one lattice keeps tiles together
    const int map = c->nrect > 0 ? 1 : 0;
    const int grid = map ? 8 * ((c->sty + 7) / 8) * c->stx : ntiles;
    const int sbegin = beside ? begin : 0;
    if (beside) {
        LBIB_HIP(hipEventRecord(c->ev_join, c->stream2));
        LBIB_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        c->topo_pending = false;
    }
    if (mixed) {
        float2* F = reinterpret_cast<float2*>(c->Fx);
        if (map)
            k_spread_fixed<1><<<grid, 64, 0, c->stream>>>(c->nx, c->ny, c->stx, c->sty, c->tcnt, c->pb_pos, c->pb_r,
                                                          c->pb_f, c->ovf, F, c->shard, sbegin, begin_it, c->st);
        else
            k_spread_fixed<0><<<grid, 64, 0, c->stream>>>(c->nx, c->ny, c->stx, c->sty, c->tcnt, c->pb_pos, c->pb_r,
                                                          c->pb_f, c->ovf, F, c->shard, sbegin, begin_it, c->st);
    } else if (map) {
        k_tile_spread<1><<<grid, SPW, 0, c->stream>>>(c->nx, c->ny, c->stx, c->sty, c->tact, c->tcnt, c->tlist, c->tpos,
                                                      c->ovf, c->mx, c->my, c->mfx, c->mfy, c->Fx, c->Fy, c->shard,
                                                      sbegin, begin_it, c->st);
    } else {
        k_tile_spread<0><<<grid, SPW, 0, c->stream>>>(c->nx, c->ny, c->stx, c->sty, c->tact, c->tcnt, c->tlist, c->tpos,
                                                      c->ovf, c->mx, c->my, c->mfx, c->mfy, c->Fx, c->Fy, c->shard,
                                                      sbegin, begin_it, c->st);
    }
    LBIB_HIP(hipGetLastError());
    return 0;
}

// Ordered bins of the markers without the forces (the junction search).
int launch_tile_bins(lbib_ctx* c) {
    launch_bin_clear(c);
    c->bins_clear = false;
    const BinArgs a = bin_args(c, false);
    k_force_bin<0><<<std::max(nblocks(c->mcap, 256), 1), 256, 0, c->stream>>>(a, 0, 0u, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_clear_force(lbib_ctx*) { return 0; }

int launch_interp_move(lbib_ctx* c) {
    if (c->nm == 0) return 0;
    const int32_t* rt = c->nrect ? c->rtile : nullptr;
    const int nt = c->stx * c->sty;
    if (c->prec == LBIB_FP64)
        k_interp_move<double><<<nblocks(c->mcap, 256), 256, 0, c->stream>>>(
            c->nx, c->ny, c->mx, c->my, c->mvx, c->mvy, c->mox, c->moy, c->u, rt, c->rects, c->stx, c->tcnt,
            c->tact, nt, c->st);
    else
        k_interp_move<float><<<nblocks(c->mcap, 256), 256, 0, c->stream>>>(
            c->nx, c->ny, c->mx, c->my, c->mvx, c->mvy, c->mox, c->moy, reinterpret_cast<const float*>(c->u), rt,
            c->rects, c->stx, c->tcnt, c->tact, nt, c->st);
    LBIB_HIP(hipGetLastError());
    c->bins_clear = true;
    return 0;
}

}  // namespace lbib
