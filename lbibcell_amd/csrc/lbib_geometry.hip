// lbib_geometry.hip -- per-step lattice "cure" on the device: connection
// checks, boundary-node (cut-link) generation with closest-wins, the local
// domain update with CDE re-initialisation, the per-row domain scanline, the
// remesh/coarsen triggers and per-cell areas.
//
// Cut links.  Connection::generateBoundaryNodes (Connection.cpp:224-324)
// visits, for each connection in vector order, every vertical lattice line it
// crosses (south node, then north node) and then every horizontal one (west,
// then east), and Connection::insertBoundaryNode (Connection.cpp:181-222)
// keeps per (site, slot) the closest node, the first inserted on ties.  Here
// each connection writes its candidates with a global insertion number
// seq = conn * 256 + k; three passes pick min distance, then min seq among
// equal distances, which is exactly the sequential outcome.  The distance
// pass marks the entries where two candidates met at the same distance; the
// sequence pass runs its atomics only there (distances rarely tie), and an
// unmarked entry's winner is its one candidate at the minimum distance.
#include "lbib_device_math.h"

namespace lbib {

__device__ __forceinline__ bool halted(const DevStatus* st) { return (st->bits & STOP_BITS) != 0u; }

// closest-wins selection entries, site-major: the four slots of a site share
// one 64-byte span
__device__ __forceinline__ long sel_index(int site, int slot) { return (long)site * 4 + (slot - 1); }
// the seq of a selection entry where two candidates met at the same distance
// (above every insertion number, below the cleared value)
constexpr uint32_t SEQ_TIE = 0xfffffffeu;

// Start of the geometry stage, one launch: the Connection::perturbConnection
// trigger (Connection.cpp:339-340, 363-364), deletion of the boundary nodes of
// the last completed cure (GeometryHandler.cpp:363-369; their candidate count
// is ncand_prev, fixed by k_cut_tie) together with the clearing of that cure's
// selection entries, and the stage counters.  A halt raised meanwhile leaves
// these to be rerun, which repeats them unchanged.
__global__ __launch_bounds__(256) void k_geo_begin(const int32_t* __restrict__ ca, const int32_t* __restrict__ cb,
                                                   const double* __restrict__ mx, const double* __restrict__ my,
                                                   const Cand* __restrict__ cand, uint32_t* tag, SelEntry* best,
                                                   uint32_t* shard, DevStatus* st) {
    if (halted(st)) return;
    const uint32_t nc = st->nc, nprev = st->ncand_prev;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += stride) {
        const int a = ca[c], b = cb[c];
        if (fabs(mx[a] - mx[b]) < kEPS || fabs(my[a] - my[b]) < kEPS)
            atomicOr(&st->bits, LBIB_HALT_PERTURB_CONN);
    }
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nprev; j += stride) {
        const uint32_t ent = cand[j].ent, site = ent >> 2;
        // every thread touching a site stores the same value: no atomic
        const uint32_t t0 = tag[site];
        if (t0 & ~DOM_MASK) tag[site] = t0 & DOM_MASK;
        best[ent] = SelEntry{~0ull, ~0u, ~0u};
    }
    if (blockIdx.x == 0 && threadIdx.x < NSHARD) shard[(SH_NBN * NSHARD + threadIdx.x) * SHARD_STRIDE] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->cure_epoch = st->cure_epoch + 1u;
        st->ncand = 0;
        st->nchanged = 0;
        st->rm_total = 0;
        st->co_removed = 0;
        st->rm_K = 0;
        st->rm_L = 0;
        st->topo_changed = 0;
        st->reinit_levels = 0;
    }
}

__device__ __forceinline__ double sqd(double x1, double y1, double x2, double y2) {
    return (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2);   // EulerianPoint.hpp:88-96
}

// Every candidate carries what its winner needs (the connection's boundary
// descriptor and, for one with a descriptor, the boundary-node velocity of
// Connection::calculateVelocity, Connection.cpp:146-168), so that no later
// kernel of the stage reads the connections or markers that the remesh
// rewrites beside it.
__global__ __launch_bounds__(1024) void k_cut_gen(int nx, int ny, long n, const int32_t* __restrict__ ca,
                          const int32_t* __restrict__ cb, const uint32_t* __restrict__ cdom,
                          const int32_t* __restrict__ cbc, const double* __restrict__ mx,
                          const double* __restrict__ my, const double* __restrict__ mvx,
                          const double* __restrict__ mvy, int want_bc, Cand* cand, CandV* cand_v,
                          int candcap, SelEntry* best, DevStatus* st) {
    if (halted(st)) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = c < (int)st->nc;
    double x1 = 0, y1 = 0, x2 = 1, y2 = 1;
    int ia = 0, ib = 0;
    if (live) { ia = ca[c]; ib = cb[c]; x1 = mx[ia]; y1 = my[ia]; x2 = mx[ib]; y2 = my[ib]; }
    const double maxX = fmax(x1, x2), maxY = fmax(y1, y2);
    const double minX = fmin(x1, x2), minY = fmin(y1, y2);
    const double m = (y2 - y1) / (x2 - x1);
    const double nn = y1 - (y2 - y1) / (x2 - x1) * x1;
    bool ok = live;
    if (live && !isfinite(m)) {
        atomicOr(&st->bits, LBIB_ERR_SLOPE);
        ok = false;
    }
    const int iv0 = (int)ceil(minX), iv1 = (int)floor(maxX);
    const int ih0 = (int)ceil(minY), ih1 = (int)floor(maxY);
    const int nv = (ok && iv1 >= iv0) ? iv1 - iv0 + 1 : 0;
    const int nh = (ok && ih1 >= ih0) ? ih1 - ih0 + 1 : 0;
    int cnt = 2 * (nv + nh);
    if (cnt > SEQ_PER_CONN) {
        atomicOr(&st->bits, LBIB_ERR_CAPACITY);
        cnt = 0;
    }
    const uint32_t base = block_alloc(&st->ncand, (uint32_t)cnt);
    if (cnt == 0) return;
    if ((long)base + cnt > candcap) {
        atomicOr(&st->bits, LBIB_ERR_CAPACITY);
        return;
    }
    const uint32_t dom = cdom[c];
    const int desc = want_bc ? cbc[c] : -1;
    double vxa = 0, vya = 0, vxb = 0, vyb = 0;
    if (desc >= 0) { vxa = mvx[ia]; vya = mvy[ia]; vxb = mvx[ib]; vyb = mvy[ib]; }
    int k = 0;
    auto put = [&](int sx, int sy, int slot, uint32_t d, double bx, double by) {
        Cand e;
        if (sx < 0 || sx >= nx || sy < 0 || sy >= ny) {
            atomicOr(&st->bits, LBIB_ERR_EDGE);
            sx = 0; sy = 0;
        }
        const long site = (long)sy * nx + sx;
        e.ent = (uint32_t)sel_index((int)site, slot);
        e.seq = (uint32_t)c * SEQ_PER_CONN + k;
        e.domain = d;
        e.dist2 = sqd((double)sx, (double)sy, bx, by);
        e.desc = desc;
        cand[base + k] = e;
        if (desc >= 0) {
            const double l = sqrt(sqd(x1, y1, bx, by));
            const double L = sqrt(sqd(x2, y2, x1, y1));
            cand_v[base + k] = CandV{(vxb - vxa) * l / L + vxa, (vyb - vya) * l / L + vya};
        }
        const unsigned long long key = (unsigned long long)__double_as_longlong(e.dist2);
        // a candidate meeting an equal distance already in place marks the
        // entry: only such entries need the insertion-number pass
        if (atomicMin(&best[e.ent].dist, key) == key) best[e.ent].seq = SEQ_TIE;
        k++;
    };
    // vertical lattice lines: node S at (i, floor y) [its slot N], node N at (i, ceil y) [slot S]
    for (int i = iv0; i < iv0 + nv; i++) {
        const double y = m * i + nn;
        const uint32_t dN = (x1 < x2) ? dom : 0u, dS = (x1 < x2) ? 0u : dom;
        put(i, (int)floor(y), DN, dS, (double)i, y);
        put(i, (int)ceil(y), DS, dN, (double)i, y);
    }
    // horizontal lattice lines: node W at (floor x, i) [slot E], node E at (ceil x, i) [slot W]
    for (int i = ih0; i < ih0 + nh; i++) {
        const double x = (i - nn) / m;
        const uint32_t dW = (y1 > y2) ? 0u : dom, dE = (y1 > y2) ? dom : 0u;
        put((int)floor(x), i, DE, dW, x, (double)i);
        put((int)ceil(x), i, DW, dE, x, (double)i);
    }
}

__global__ void k_cut_tie(const Cand* __restrict__ cand, long n, SelEntry* best, DevStatus* st) {
    if (halted_beside_topology(st)) return;
    const uint32_t nc = st->ncand;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->ncand_prev = nc;   // this cure's reset list
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x) {
        const Cand e = cand[j];
        SelEntry& b = best[e.ent];
        if (b.seq != ~0u && (unsigned long long)__double_as_longlong(e.dist2) == b.dist) atomicMin(&b.seq, e.seq);
    }
}

// rmark[site] of a site listed for re-initialisation: the cure that listed it
// (a counter k_geo_begin advances) and its position p in the changed list, so
// that entries left by earlier cures never match.
__device__ __forceinline__ uint64_t rmark_val(uint32_t epoch, uint32_t p) { return ((uint64_t)epoch << 32) | p; }

// Winners: mark the slot, record the domain and (for BC-carrying connections)
// the node velocity of Connection::calculateVelocity (Connection.cpp:146-168).
// Then PhysicalNode::updateDomainIdentifier (PhysicalNode.cpp:109-124): a site
// takes the domain of its last non-null boundary neighbour in slot order
// E,N,W,S; every slot with a candidate has exactly one winner, so that is the
// winner in the site's highest slot with a candidate (a set best distance),
// which reports the site when its domain changes.
// A workgroup takes chunks of WIN_PER_WG candidates (two per lane) and lists
// its reported sites with one atomic per chunk; the winners are counted in
// sharded counters.
constexpr int WIN_WG = 1024, WIN_PER_WG = 2 * WIN_WG;
struct WinOut {
    uint32_t won, add, site, dom;
};
__device__ __forceinline__ WinOut cut_winner(const Cand* cand, const CandV* cand_v, uint32_t j, uint32_t nc, long n,
                                             const SelEntry* __restrict__ best, uint32_t* tag, uint32_t* bn_dom,
                                             double* bc_v, int32_t* bc_desc, int local_update) {
    WinOut o{0u, 0u, 0u, 0u};
    if (j >= nc) return o;
    const Cand e = cand[j];
    const uint32_t site = e.ent >> 2;
    const int slot = (int)(e.ent & 3u) + 1;
    const long slotix = (long)(slot - 1) * n + site;
    const uint32_t dom_now = tag[site] & DOM_MASK;   // domain bits are not changed here
    const SelEntry b = best[e.ent];
    if (b.seq == ~0u ? (unsigned long long)__double_as_longlong(e.dist2) == b.dist : b.seq == e.seq) {
        o.won = 1;
        o.site = site;
        o.dom = e.domain;
        bn_dom[slotix] = e.domain;
        uint32_t bits = 1u << (BN_SHIFT + slot - 1);
        if (e.desc >= 0 && bc_v != nullptr) {
            bits |= 1u << (BC_SHIFT + slot - 1);
            const CandV v = cand_v[j];
            bc_v[2 * slotix] = v.vx;
            bc_v[2 * slotix + 1] = v.vy;
            bc_desc[slotix] = e.desc;
        }
        atomicOr(&tag[site], bits);
        if (local_update && dom_now != e.domain) {
            bool highest = true;
            for (int s2 = slot + 1; s2 <= 4; s2++)
                if (best[sel_index((int)site, s2)].dist != ~0ull) highest = false;
            o.add = highest ? 1u : 0u;
        }
    }
    return o;
}

__global__ __launch_bounds__(WIN_WG) void k_cut_win(const Cand* cand, const CandV* cand_v, long n,
                                                    const SelEntry* __restrict__ best, uint32_t* tag,
                                                    uint32_t* bn_dom, double* bc_v, int32_t* bc_desc,
                                                    uint32_t* changed, uint64_t* rmark, uint8_t* rstate,
                                                    uint32_t* shard, int local_update, DevStatus* st) {
    if (halted_beside_topology(st)) return;
    const uint32_t nc = st->ncand, epoch = st->cure_epoch;
    for (uint32_t j0 = blockIdx.x * (uint32_t)WIN_PER_WG; j0 < nc; j0 += gridDim.x * (uint32_t)WIN_PER_WG) {
        // both candidates' loads in flight together
        const WinOut a = cut_winner(cand, cand_v, j0 + threadIdx.x, nc, n, best, tag, bn_dom, bc_v, bc_desc,
                                    local_update);
        const WinOut b = cut_winner(cand, cand_v, j0 + WIN_WG + threadIdx.x, nc, n, best, tag, bn_dom, bc_v, bc_desc,
                                    local_update);
        const uint32_t nw = (uint32_t)__popcll(__ballot(a.won != 0u)) + (uint32_t)__popcll(__ballot(b.won != 0u));
        if ((threadIdx.x & 63) == 0 && nw)
            atomicAdd(&shard[(SH_NBN * NSHARD + ((blockIdx.x * (WIN_WG / 64) + threadIdx.x / 64) % NSHARD)) *
                             SHARD_STRIDE],
                      nw);
        const uint32_t p0 = block_alloc(&st->nchanged, a.add + b.add);
        auto report = [&](const WinOut& w, uint32_t p) {
            if (p < (uint32_t)REINIT_CAP) {
                changed[2 * p] = w.site;
                changed[2 * p + 1] = w.dom;
                rmark[w.site] = rmark_val(epoch, p);
                rstate[p] = 0;
            } else {
                atomicOr(&st->bits, LBIB_ERR_CAPACITY);
            }
        };
        if (a.add) report(a, p0);
        if (b.add) report(b, p0 + a.add);
    }
}

struct ReinitArgs {
    int nx, ny;
    long n;
    int ns;
    void* g;             // current species buffer [ns][5][n] in the storage format
    const void* u;       // velocity pairs (ux, uy): FP64, or FP32 for mixed precision
    uint32_t* tag;
    const uint32_t* changed;
    const uint64_t* rmark;
    uint32_t it;
    int stx;
    const DevStatus* st;    // iteration of ITER_NEXT
    const int32_t* rtile;   // batched atlas (null: one periodic domain)
    const RectDev* rects;
    SpeciesDev sp[LBIB_MAX_SPECIES];
};

// P: storage format; the re-initialisation itself is FP64 arithmetic.  The
// neighbour tags and the axis neighbours' concentrations are loaded before
// any is used (the sums below keep the reference's order), so a site costs a
// few dependent memory round trips rather than one per neighbour.
template <int NS, int P>
__device__ void reinit_site(const ReinitArgs& a, const long i, const uint32_t nd) {
    const int x = (int)(i % a.nx), y = (int)(i / a.nx);
    const int EXd[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    const int EYd[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    Frame fr;
    if (!site_frame(a.rtile, a.rects, a.stx, a.nx, a.ny, x, y, fr)) {
        fr.x0 = 0; fr.y0 = 0; fr.w = a.nx; fr.h = a.ny; fr.r = -1;   // no domain owns it
    }
    long nbi[9];
#pragma unroll
    for (int d = 1; d < 9; d++) {
        const int xx = fr.x0 + (x - fr.x0 + EXd[d] + fr.w) % fr.w;
        const int yy = fr.y0 + (y - fr.y0 + EYd[d] + fr.h) % fr.h;
        nbi[d] = (long)yy * a.nx + xx;
    }
    constexpr int GQ = NS > 0 ? 5 * NS : 5;   // species planes in the storage blocks
    using GS = GStore<P>;
    auto cswap = [&](int k, long j) -> double {
        return c_swapped(GS::get(a.g, k, 0, j, a.n, GQ), GS::get(a.g, k, 1, j, a.n, GQ), GS::get(a.g, k, 2, j, a.n, GQ),
                         GS::get(a.g, k, 3, j, a.n, GQ), GS::get(a.g, k, 4, j, a.n, GQ));
    };
    // the neighbours' domains (no neighbour is re-initialised concurrently:
    // sites of one level are never adjacent) and the site's own values
    uint32_t tn[9];
#pragma unroll
    for (int d = 1; d < 9; d++) tn[d] = a.tag[nbi[d]] & DOM_MASK;
    const uint32_t t0 = a.tag[i];
    double ux_i, uy_i;
    if (P == LBIB_FP64) {
        ux_i = static_cast<const double*>(a.u)[2 * i];
        uy_i = static_cast<const double*>(a.u)[2 * i + 1];
    } else {
        ux_i = (double)static_cast<const float*>(a.u)[2 * i];
        uy_i = (double)static_cast<const float*>(a.u)[2 * i + 1];
    }
    const uint32_t tg = (t0 & ~DOM_MASK) | nd;
    a.tag[i] = tg;
    const int dir1[4] = {DE, DN, DW, DS};
    const int dir2[4] = {DNE, DNW, DSW, DSE};
    for (int k = 0; k < NS; k++) {
        double c1[4];
#pragma unroll
        for (int q = 0; q < 4; q++) c1[q] = cswap(k, nbi[dir1[q]]);
        double sumC = 0.0;
        int counter = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int d = dir1[q];
            const bool hasbn = (tg >> (BN_SHIFT + d - 1)) & 1u;
            if (!hasbn && nd == tn[d]) { sumC += c1[q]; counter++; }
        }
        if (counter == 0) {
            for (int q = 0; q < 4; q++)
                if (nd == tn[dir2[q]]) { sumC += cswap(k, nbi[dir2[q]]); counter++; }
        }
        if (counter == 0 && a.sp[k].kind != LBIB_SPECIES_PLAIN) {
#pragma unroll
            for (int q = 0; q < 4; q++) { sumC += c1[q]; counter++; }
        }
        sumC /= static_cast<double>(counter);
        double d[5];
        for (int q = 0; q < 5; q++) d[q] = sumC / 5.0;
        double C = 0.0;
        for (int q = 0; q < 5; q++) C += d[q];
        double cj[NS > 0 ? NS : 1];
        for (int j = 0; j < NS; j++) cj[j] = (j == k) ? C : cswap(j, i);
        const double tI = fr.r < 0 ? a.sp[k].tauI : a.rects[fr.r].tauIs[k];
        species_collide<NS>(d, C, ux_i, uy_i, a.sp[k], tI, nd, iter_of(a.it, a.st), cj, cj);
        GS::put5(a.g, k, i, a.n, GQ, d);
    }
}

// GeometryHandler::updateDomainIdentifier (GeometryHandler.cpp:58-66) runs the
// per-site update in (y, x) order and each re-initialisation
// (CDESolverD2Q5.cpp:156-228, ..._SIGNAL.cpp:184-250) reads neighbours updated
// earlier in that order.  The sequential pass is replayed level by level: a
// changed site is ready once every changed neighbour that precedes it in
// (y, x) order (a smaller site index) has been processed; neighbours that
// follow it still hold their old values when it runs, as in the reference.
// Ready sites are never adjacent, so each level runs in parallel and the
// result is identical.  A changed neighbour is found through rmark[site]
// (its position in the list, tagged with the cure epoch).
//
// The first REINIT_GRID_LEVELS levels run as one grid-wide launch each (most
// sites are ready at once); state 2 + L marks a site done by level launch L,
// so a launch only trusts neighbours finished by earlier launches.  The rest
// runs in one workgroup, level after level with barriers (state 1 = ready,
// >= 2 = done).
constexpr int REINIT_GRID_LEVELS = 2;

// Are all changed neighbours preceding site i finished (state in [lo, hi))?
// The eight neighbours' marks are loaded at once, then the states of the
// listed ones.  A changed neighbour carries this cure's epoch in its mark.
__device__ __forceinline__ bool reinit_ready(const ReinitArgs& a, uint32_t cnt, uint32_t epoch, uint32_t i,
                                             const uint8_t* state, int lo, int hi) {
    const int EXd[8] = {1, 0, -1, 0, 1, -1, -1, 1};
    const int EYd[8] = {0, 1, 0, -1, 1, 1, -1, -1};
    const int x = (int)(i % a.nx), y = (int)(i / a.nx);
    uint64_t m[8];
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const int xx = (x + EXd[d] + a.nx) % a.nx, yy = (y + EYd[d] + a.ny) % a.ny;
        const uint32_t j = (uint32_t)yy * a.nx + xx;
        m[d] = j < i ? a.rmark[j] : 0ull;
    }
    bool ready = true;
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const uint32_t p = (uint32_t)m[d];
        if ((uint32_t)(m[d] >> 32) == epoch && p < cnt) {
            if (hi <= lo) return false;   // a listed predecessor: not ready at this level
            const int s = state[p];
            ready = ready && s >= lo && s < hi;
        }
    }
    return ready;
}

template <int NS, int P>
__global__ __launch_bounds__(256) void k_reinit_level(const ReinitArgs a, uint8_t* state, int level, DevStatus* st) {
    if (halted_beside_topology(st)) return;
    const uint32_t cnt = min(st->nchanged, (uint32_t)REINIT_CAP), epoch = st->cure_epoch;
    bool any = false;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += gridDim.x * blockDim.x) {
        if (state[t] != 0) continue;
        const uint32_t i = a.changed[2 * t];
        if (!reinit_ready(a, cnt, epoch, i, state, 2, 2 + level)) continue;
        reinit_site<NS, P>(a, (long)i, a.changed[2 * t + 1]);
        state[t] = (uint8_t)(2 + level);
        any = true;
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) atomicMax(&st->reinit_levels, (uint32_t)level + 1);
}

template <int NS, int P>
__global__ __launch_bounds__(1024) void k_reinit(const ReinitArgs a, uint8_t* gstate, DevStatus* st) {
    if (halted_beside_topology(st)) return;
    __shared__ uint8_t lstate[REINIT_LDS];
    __shared__ int pending;
    const uint32_t cnt = min(st->nchanged, (uint32_t)REINIT_CAP), epoch = st->cure_epoch;
    if (cnt == 0) return;
    const bool in_lds = cnt <= (uint32_t)REINIT_LDS;
    uint8_t* state = in_lds ? lstate : gstate;
    if (in_lds)
        for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) lstate[t] = gstate[t];
    __syncthreads();
    for (int level = 0; level < (1 << 20); level++) {
        if (threadIdx.x == 0) pending = 0;
        __syncthreads();
        // phase 1: mark the ready sites (state 1)
        bool left = false;
        for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
            if (state[t] != 0) continue;
            left = true;
            if (reinit_ready(a, cnt, epoch, a.changed[2 * t], state, 2, 256)) state[t] = 1;
        }
        if (left) pending = 1;
        __syncthreads();
        if (!pending) {
            if (threadIdx.x == 0 && level > 0) st->reinit_levels = REINIT_GRID_LEVELS + (uint32_t)level;
            break;
        }
        // phase 2: run them
        for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
            if (state[t] != 1) continue;
            reinit_site<NS, P>(a, (long)a.changed[2 * t], a.changed[2 * t + 1]);
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x)
            if (state[t] == 1) state[t] = 2;
        __syncthreads();
    }
}

// GeometryHandler::updateAllDomainIdentifiers (GeometryHandler.cpp:68-85): per
// row, the domain of the last west boundary node at or left of x (x >= 1), or
// the domain of site (0, y), which must be 0.
// Each thread owns 8 consecutive sites of a 2048-site chunk; the last west
// boundary node before a thread's sites comes from a wavefront max-scan of
// positions and the wavefronts before it.
__global__ __launch_bounds__(256) void k_scanline(int nx, long n, uint32_t* tag, const uint32_t* __restrict__ bn_dom,
                                                  DevStatus* st) {
    if (halted_beside_topology(st)) return;
    constexpr int PER = 8;
    __shared__ int wpos[4];
    __shared__ uint32_t wdom[4];
    const int y = blockIdx.x;
    const long r0 = (long)y * nx;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry = tag[r0] & DOM_MASK;
    if (threadIdx.x == 0 && carry != 0u) atomicOr(&st->bits, LBIB_ERR_SCANLINE);
    const int wbit = BN_SHIFT + DW - 1;
    const uint32_t* wdom_row = bn_dom + (long)(DW - 1) * n + r0;
    for (int x0 = 0; x0 < nx; x0 += 256 * PER) {
        const int xb = x0 + (int)threadIdx.x * PER;
        uint32_t tg[PER], dm[PER];
        int lpos = -1;
        uint32_t ldom = 0;
        if (xb + PER <= nx && ((r0 + xb) & 3) == 0) {
            const uint4* p4 = reinterpret_cast<const uint4*>(tag + r0 + xb);
            const uint4 u0 = p4[0], u1 = p4[1];
            tg[0] = u0.x; tg[1] = u0.y; tg[2] = u0.z; tg[3] = u0.w;
            tg[4] = u1.x; tg[5] = u1.y; tg[6] = u1.z; tg[7] = u1.w;
        } else {
#pragma unroll
            for (int q = 0; q < PER; q++) tg[q] = (xb + q < nx) ? tag[r0 + xb + q] : 0u;
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int x = xb + q;
            dm[q] = 0;
            if (x >= 1 && x < nx && ((tg[q] >> wbit) & 1u)) {
                dm[q] = wdom_row[x];
                lpos = x;
                ldom = dm[q];
            }
        }
        // inclusive max-scan of (position, domain) over the wavefront
        int p = lpos;
        uint32_t d = ldom;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int p2 = __shfl_up(p, o, 64);
            const uint32_t d2 = __shfl_up(d, o, 64);
            if (lane >= o && p2 > p) { p = p2; d = d2; }
        }
        int pe = __shfl_up(p, 1, 64);
        uint32_t de = __shfl_up(d, 1, 64);
        if (lane == 0) pe = -1;
        if (lane == 63) { wpos[w] = p; wdom[w] = d; }
        __syncthreads();
        int bp = -1;
        uint32_t cur = carry;
        for (int k = 0; k < w; k++)
            if (wpos[k] > bp) { bp = wpos[k]; cur = wdom[k]; }
        if (pe > bp) cur = de;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int x = xb + q;
            if (x >= 1 && x < nx) {
                if ((tg[q] >> wbit) & 1u) cur = dm[q];
                const uint32_t nt = (tg[q] & ~DOM_MASK) | cur;
                if (nt != tg[q]) tag[r0 + x] = nt;
            }
        }
        for (int k = 0; k < 4; k++)
            if (wpos[k] >= 0) carry = wdom[k];
        __syncthreads();
    }
}

template <int NS, int P>
static void launch_reinit_p(lbib_ctx* c, const ReinitArgs& a) {
    for (int level = 0; level < REINIT_GRID_LEVELS; level++)
        k_reinit_level<NS, P><<<64, 256, 0, c->stream>>>(a, c->rstate, level, c->st);
    k_reinit<NS, P><<<1, 1024, 0, c->stream>>>(a, c->rstate, c->st);
}
template <int NS>
static void launch_reinit_t(lbib_ctx* c, const ReinitArgs& a) {
    switch (c->prec) {
        case LBIB_FP64: launch_reinit_p<NS, LBIB_FP64>(c, a); break;
        case LBIB_FP32: launch_reinit_p<NS, LBIB_FP32>(c, a); break;
        case LBIB_FP16_DEV: launch_reinit_p<NS, LBIB_FP16_DEV>(c, a); break;
        case LBIB_FP16_REM: launch_reinit_p<NS, LBIB_FP16_REM>(c, a); break;
        default: launch_reinit_p<NS, LBIB_BF16_DEV>(c, a); break;
    }
}

int launch_topology(lbib_ctx* c, bool coarsen);

int launch_geometry(lbib_ctx* c, uint32_t it, int mode) {
    // mode: 0 = moveLattice (local update + reinit), 1 = cureLattice (scanline only),
    // bit 2 = check coarsening
    const bool cure = (mode & 1) != 0;
    const bool coarsen = (mode & 2) != 0;
    const int bs = 256;
    const int gs = 1024;
    // connection counts are device-resident (remeshing changes them): grids
    // cover the capacity and the kernels read st->nc
    k_geo_begin<<<gs, bs, 0, c->stream>>>(c->ca, c->cb, c->mx, c->my, c->cand, c->tag, c->best, c->shard, c->st);
    // 1024-wide workgroups: one candidate-counter atomic per 1024 connections
    if (c->ccap > 0)
        k_cut_gen<<<(c->ccap + 1023) / 1024, 1024, 0, c->stream>>>(c->nx, c->ny, c->n, c->ca, c->cb, c->cdom, c->cbc,
                                                               c->mx, c->my, c->mvx, c->mvy, c->bc_v ? 1 : 0,
                                                               c->cand, c->cand_v, c->candcap, c->best, c->st);
    // From here on nothing of this stage reads the connections or markers:
    // the topology stage (remesh, rings, coarsening), which reads and
    // rewrites them, runs beside the rest on the second stream.
    const bool beside = !cure && c->nc > 0;
    if (beside) {
        LBIB_HIP(hipEventRecord(c->ev_fork, c->stream));
        LBIB_HIP(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
        hipStream_t main = c->stream;
        c->stream = c->stream2;
        const int r = launch_topology(c, coarsen);
        c->stream = main;
        if (r) return -1;
    }
    k_cut_tie<<<gs, bs, 0, c->stream>>>(c->cand, c->n, c->best, c->st);
    // enough workgroups for about twice the connections' candidates (the
    // count is device-resident; the kernel loops over more)
    const int wgrid = (int)std::min<long>(std::max<long>((2L * std::max(c->nc, 1) + WIN_PER_WG - 1) / WIN_PER_WG, 64),
                                          1024);
    k_cut_win<<<wgrid, WIN_WG, 0, c->stream>>>(c->cand, c->cand_v, c->n, c->best, c->tag, c->bn_dom, c->bc_v,
                                               c->bc_desc, c->changed, c->rmark, c->rstate, c->shard, cure ? 0 : 1,
                                               c->st);
    if (!cure) {
        ReinitArgs a;
        a.nx = c->nx; a.ny = c->ny; a.n = c->n; a.ns = c->ns;
        a.g = c->g[c->cur];
        a.u = c->u; a.tag = c->tag; a.changed = c->changed; a.rmark = c->rmark; a.it = it;
        a.stx = c->stx; a.rtile = c->nrect ? c->rtile : nullptr; a.rects = c->rects; a.st = c->st;
        for (int s = 0; s < LBIB_MAX_SPECIES; s++) a.sp[s] = c->sp[s];
        switch (c->ns) {
            case 0: launch_reinit_t<0>(c, a); break;
            case 1: launch_reinit_t<1>(c, a); break;
            case 2: launch_reinit_t<2>(c, a); break;
            case 3: launch_reinit_t<3>(c, a); break;
            default: launch_reinit_t<4>(c, a); break;
        }
    }
    k_scanline<<<c->ny, 256, 0, c->stream>>>(c->nx, c->n, c->tag, c->bn_dom, c->st);
    LBIB_HIP(hipGetLastError());
    // the main stream waits for the topology stage at the next entry point
    // (join_topology), or after the next lattice stage's marker forces and
    // binning, which then follow the topology stage on the second stream
    if (beside) {
        c->topo_pending = true;
        c->forces_beside = true;
    }
    return 0;
}

// remeshBoundary, then at it % 10 == 0 deleteAllForces + coarsenBoundary
// (SimulationRunner.cpp:83-87); rings are rebuilt for the area helpers.
int launch_topology(lbib_ctx* c, bool coarsen) {
    if (c->nc == 0) return 0;
    if (launch_remesh(c, 1)) return -1;
    // the ring order is rebuilt only when remeshing or coarsening changed
    // the topology in this step (st->topo_changed)
    if (launch_rings(c, true, true)) return -1;
    if (coarsen) {
        if (launch_coarsen(c)) return -1;
        if (launch_rings(c, false, true)) return -1;
    }
    return 0;
}

int launch_area_regulator(lbib_ctx* c, double target, double kp, double maxs, double factor) {
    if (c->ndom == 0) return 0;
    const AreaMass am{c->celltype, c->target, c->mtab, target, kp, maxs, factor};
    if (launch_ring_areas(c, &am)) return -1;
    c->mtab_live = true;
    return 0;
}

}  // namespace lbib
