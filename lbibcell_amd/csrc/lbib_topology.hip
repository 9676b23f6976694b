// lbib_topology.hip -- membrane topology kept on the device: remeshing,
// coarsening, ring order for the cell areas, marker in/out links and the
// membrane-tension force table.
//
// Remesh (GeometryHandler::remeshBoundary, GeometryHandler.cpp:385-431): every
// connection longer than MAXLENGTH is erased from connections_ and its two
// halves are appended in visiting order, with new GeometryNode ids max+1,
// max+2, ... .  One round = flag, exclusive scan, scatter: survivors keep
// their relative order, the halves of the r-th long connection land at
// (nc - L) + 2r and (nc - L) + 2r + 1.
//
// Coarsen (GeometryHandler::coarsenBoundary, :433-461): markers are visited in
// id order; removing one splices its neighbours with a new connection appended
// to the vector.  Rings do not share markers, so each ring is walked by one
// lane in id order inside LDS; the appended connections are then ordered by
// the id of the marker whose removal created them (the global visiting order),
// and those erased again by a later removal are dropped.
//
// Ring order (GeometryHandler::computeAreas, :101-134): the polygon of a domain
// starts at its first connection in vector order and follows the outgoing
// links; positions come from pointer jumping in LDS.
#include <algorithm>
#include <climits>
#include "lbib_device_math.h"

namespace lbib {

__device__ __forceinline__ bool topo_halted(const DevStatus* st) { return (st->bits & STOP_BITS) != 0u; }

// ------------------------------------------------------------------ scan ----
// Exclusive scan of n (device-resident count) uint32 values, blocks of 1024.
constexpr int SCAN_B = 1024;

// gate != nullptr: *gate == 0 skips the whole scan (outputs and total untouched)
__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t* __restrict__ in, const uint32_t* nptr, uint32_t nfix,
                                                     uint32_t* __restrict__ bsum, const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    const uint32_t n = nptr ? *nptr : nfix;
    const uint32_t b0 = blockIdx.x * SCAN_B;
    __shared__ uint32_t wsum[16];
    uint32_t s = 0;
    if (b0 < n)
        for (uint32_t k = threadIdx.x; k < SCAN_B; k += 256) {
            const uint32_t i = b0 + k;
            if (i < n) s += in[i];
        }
    uint32_t ex;
    const uint32_t total = block_excl_scan(s, &ex, wsum);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// single block: exclusive scan of the block sums; total -> *total
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* bsum, uint32_t nb, uint32_t* total, const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    __shared__ uint32_t wsum[16];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = (i < nb) ? bsum[i] : 0u;
        uint32_t ex;
        const uint32_t t = block_excl_scan(v, &ex, wsum);
        if (i < nb) bsum[i] = carry + ex;
        carry += t;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* __restrict__ in, const uint32_t* nptr, uint32_t nfix,
                                                    const uint32_t* __restrict__ bsum, uint32_t* __restrict__ out,
                                                    const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    const uint32_t n = nptr ? *nptr : nfix;
    const uint32_t b0 = blockIdx.x * SCAN_B;
    if (b0 >= n) return;
    __shared__ uint32_t wsum[16];
    // each thread owns 4 consecutive elements
    uint32_t v[4];
    uint32_t loc = 0;
    for (int q = 0; q < 4; q++) {
        const uint32_t i = b0 + threadIdx.x * 4 + q;
        v[q] = (i < n) ? in[i] : 0u;
        loc += v[q];
    }
    uint32_t ex;
    (void)block_excl_scan(loc, &ex, wsum);
    uint32_t run = bsum[blockIdx.x] + ex;
    for (int q = 0; q < 4; q++) {
        const uint32_t i = b0 + threadIdx.x * 4 + q;
        if (i < n) out[i] = run;
        run += v[q];
    }
}

// Short arrays (tiles, domains): one workgroup, 4096 values per pass.
constexpr uint32_t SCAN_SINGLE_MAX = 1u << 16;
// nz != nullptr: also count the nonzero bytes of nz[0..n) into *nz_total.
// A pass covers 16384 values: coalesced loads into LDS, each thread then scans
// 16 consecutive ones (row stride 17 words avoids bank conflicts), coalesced
// stores from LDS.
// Called by a whole 1024-wide workgroup.
__device__ void scan_single_dev(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out,
                                uint32_t* total, const uint8_t* __restrict__ nz, uint32_t* nz_total) {
    constexpr int PER = 16, PASS = 1024 * PER, ROW = PER + 1;
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t buf[1024 * ROW];
    uint32_t carry = 0, nzc = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += PASS) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const uint32_t e = (uint32_t)q * 1024u + threadIdx.x;   // element within the pass
            const uint32_t i = b0 + e;
            const uint32_t v = (i < n) ? in[i] : 0u;
            if (nz && i < n && nz[i]) nzc++;
            buf[(e / PER) * ROW + e % PER] = v;
        }
        __syncthreads();
        uint32_t* row = buf + threadIdx.x * ROW;
        uint32_t loc = 0;
#pragma unroll
        for (int q = 0; q < PER; q++) loc += row[q];
        uint32_t ex;
        const uint32_t t = block_excl_scan(loc, &ex, wsum);
        uint32_t run = carry + ex;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const uint32_t v = row[q];
            row[q] = run;
            run += v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const uint32_t e = (uint32_t)q * 1024u + threadIdx.x;
            const uint32_t i = b0 + e;
            if (i < n) out[i] = buf[(e / PER) * ROW + e % PER];
        }
        carry += t;
        __syncthreads();
    }
    if (nz) {
        uint32_t ex;
        const uint32_t t = block_excl_scan(nzc, &ex, wsum);
        if (threadIdx.x == 0) *nz_total = t;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(1024) void k_scan_single(const uint32_t* __restrict__ in, const uint32_t* nptr,
                                                      uint32_t nfix, uint32_t* __restrict__ out, uint32_t* total,
                                                      const uint8_t* __restrict__ nz, uint32_t* nz_total,
                                                      const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    scan_single_dev(in, nptr ? *nptr : nfix, out, total, nz, nz_total);
}

int scan_exclusive(lbib_ctx* c, const uint32_t* in, uint32_t* out, const uint32_t* nptr, uint32_t nfix, uint32_t cap,
                   uint32_t* total, const uint32_t* gate) {
    const uint32_t nb = (cap + SCAN_B - 1) / SCAN_B;
    if (nb == 0) return 0;
    if (cap <= SCAN_SINGLE_MAX) {
        k_scan_single<<<1, 1024, 0, c->stream>>>(in, nptr, nfix, out, total, nullptr, nullptr, gate);
        LBIB_HIP(hipGetLastError());
        return 0;
    }
    if (nb > c->scan_cap) {
        c->err = "scan scratch too small";
        return -1;
    }
    k_scan_blocks<<<nb, 256, 0, c->stream>>>(in, nptr, nfix, c->scan_tmp, gate);
    k_scan_top<<<1, 1024, 0, c->stream>>>(c->scan_tmp, nb, total, gate);
    k_scan_apply<<<nb, 256, 0, c->stream>>>(in, nptr, nfix, c->scan_tmp, out, gate);
    LBIB_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------- in/out ------
__global__ void k_inout(const int32_t* __restrict__ ca, const int32_t* __restrict__ cb, int32_t* min_c,
                        int32_t* mout_c, const DevStatus* st) {
    const uint32_t nc = st->nc;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x) {
        mout_c[ca[j]] = (int32_t)j;
        min_c[cb[j]] = (int32_t)j;
    }
}

// ------------------------------------------------------------- remesh ------
// Perturbation trigger of the GeometryNode ctor (GeometryNode.cpp:153-214)
__device__ __forceinline__ bool needs_perturbation(const double X, const double Y) {
    const bool p1 = fabs((double)(int)X - X) <= kEPS;
    const bool p2 = fabs((double)(int)Y - Y) <= kEPS;
    const bool p3 = fabs(Y - floor(Y) - (X - floor(X))) <= kEPS;
    const bool p4 = fabs(ceil(Y) - Y - (X - floor(X))) <= kEPS;
    return p1 || p2 || p3 || p4;
}

// flag != nullptr: mark this round's long connections (and the perturbation
// halt of their midpoints) and list them, in any order, in list (st->rm_K
// counts them; the first RM_LIST_CAP are stored);
// flag == nullptr: only raise LBIB_FLAG_REMESH if a connection is still long
// after the rounds of this pass.  ring_start != nullptr: also clear the
// per-domain ring counters for the ring rebuild that follows, when the
// topology changed (like k_ring_reset).
__global__ void k_rm_flag(const int32_t* __restrict__ ca, const int32_t* __restrict__ cb, const double* __restrict__ mx,
                          const double* __restrict__ my, uint32_t* flag, uint32_t* list, int ndom, int32_t* ring_start,
                          uint32_t* ring_cnt, DevStatus* st) {
    if (ring_start && st->topo_changed)
        for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < ndom; d += gridDim.x * blockDim.x) {
            ring_start[d] = INT_MAX;
            ring_cnt[d] = 0;
        }
    if (topo_halted(st)) return;
    const uint32_t nc = st->nc;
    bool any = false;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const int a = ca[c], b = cb[c];
        const double len = sqrt((mx[a] - mx[b]) * (mx[a] - mx[b]) + (my[a] - my[b]) * (my[a] - my[b]));
        const bool lg = len > kMAXLENGTH;
        any |= lg;
        if (!flag) continue;
        flag[c] = lg ? 1u : 0u;
        // one list slot range per wavefront: the lowest long lane takes it
        const unsigned long long m = __ballot(lg);
        if (lg) {
            const int lane = threadIdx.x & 63;
            const int lead = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == lead) base = atomicAdd(&st->rm_K, (uint32_t)__popcll(m));
            base = (uint32_t)__shfl((int)base, lead, 64);
            const uint32_t p = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (p < RM_LIST_CAP) list[p] = c;
            if (needs_perturbation((mx[a] + mx[b]) / 2, (my[a] + my[b]) / 2))
                atomicOr(&st->bits, LBIB_HALT_PERTURB_REMESH);
        }
    }
    if (!flag && __any(any) && (threadIdx.x & 63) == 0) atomicOr(&st->bits, LBIB_FLAG_REMESH);
}

// One workgroup of 1024: the round's ranks.  Up to RM_LIST_CAP long
// connections: their list sorted ascending (a connection's rank is then the
// number of entries below its index); more: an exclusive scan of the flags
// into rank.  Sets rm_L and rm_sparse and empties the list count for the next
// round, also when the topology stage is stopped.
__global__ __launch_bounds__(1024) void k_rm_rank(const uint32_t* __restrict__ flag, uint32_t* rank, uint32_t* list,
                                                  DevStatus* st) {
    __shared__ uint32_t key[RM_LIST_CAP];
    const uint32_t K = st->rm_K;
    const bool run = !topo_halted(st);
    if (run && K > 0 && K <= RM_LIST_CAP) {
        uint32_t P = 1;
        while (P < K) P <<= 1;
        for (uint32_t j = threadIdx.x; j < P; j += blockDim.x) key[j] = j < K ? list[j] : 0xffffffffu;
        __syncthreads();
        // bitonic sort of P keys (indices are distinct)
        for (uint32_t size = 2; size <= P; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t j = threadIdx.x; j < P; j += blockDim.x) {
                    const uint32_t o = j ^ stride;
                    if (o > j) {
                        const uint32_t u = key[j], v = key[o];
                        if ((u > v) == ((j & size) == 0)) { key[j] = v; key[o] = u; }
                    }
                }
                __syncthreads();
            }
        for (uint32_t j = threadIdx.x; j < K; j += blockDim.x) list[j] = key[j];
    } else if (run && K > RM_LIST_CAP) {
        scan_single_dev(flag, st->nc, rank, nullptr, nullptr, nullptr);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (run) {
            st->rm_L = K;
            st->rm_sparse = K <= RM_LIST_CAP ? 1u : 0u;
        }
        st->rm_K = 0;
    }
}

struct TopoArrays {
    double *mx, *my, *mvx, *mvy, *mox, *moy;
    uint32_t* mid;
    int32_t *ca, *cb, *cbc;
    uint32_t* cdom;
    int32_t *ca2, *cb2, *cbc2;
    uint32_t* cdom2;
    int32_t* fs_off;
};

// Scatter one remesh round into the scratch connection arrays and append the
// midpoint markers.  L = number of long connections (st->rm_L).  Capacity is
// checked first, before anything is written (every block evaluates the same
// condition; one raises the halt).
__global__ __launch_bounds__(256) void k_rm_apply(TopoArrays t, const uint32_t* __restrict__ flag,
                                                  const uint32_t* __restrict__ rank, const uint32_t* __restrict__ list,
                                                  int nx, int ny, uint32_t mcap, uint32_t ccap, DevStatus* st) {
    __shared__ uint32_t key[RM_LIST_CAP];
    __shared__ int go;
    // one reading of the status per workgroup, so that all its threads agree
    if (threadIdx.x == 0) go = !topo_halted(st) && st->rm_L > 0;
    __syncthreads();
    if (!go) return;
    const uint32_t L = st->rm_L;
    const uint32_t nc = st->nc, nm = st->nm, maxid = st->maxid;
    if (nm + L + 1 > mcap || nc + L > ccap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&st->bits, LBIB_HALT_GROW);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rm_nc0 = nc;
        st->rm_nm0 = nm;
    }
    const bool sparse = st->rm_sparse != 0u;   // then L <= RM_LIST_CAP
    if (sparse) {
        for (uint32_t j = threadIdx.x; j < L; j += blockDim.x) key[j] = list[j];
        __syncthreads();
    }
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        uint32_t r;
        if (sparse) {   // entries of the sorted list below c
            uint32_t lo = 0, hi = L;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (key[mid] < c) lo = mid + 1;
                else hi = mid;
            }
            r = lo;
        } else {
            r = rank[c];
        }
        if (!flag[c]) {
            const uint32_t p = c - r;
            t.ca2[p] = t.ca[c]; t.cb2[p] = t.cb[c]; t.cdom2[p] = t.cdom[c]; t.cbc2[p] = t.cbc[c];
            continue;
        }
        const int a = t.ca[c], b = t.cb[c];
        const uint32_t k = nm + r;
        const double x = (t.mx[a] + t.mx[b]) / 2, y = (t.my[a] + t.my[b]) / 2;
        t.mx[k] = x; t.my[k] = y;
        t.mvx[k] = 0.0; t.mvy[k] = 0.0;
        t.mox[k] = x - 5.0; t.moy[k] = y - 5.0;   // GeometryNode.cpp:102
        t.mid[k] = maxid + 1 + r;
        // the new node is linked at once (update needed): boundary checks of
        // GeometryHandler.cpp:613-636
        if (x - 2.0 < 0 || y - 2.0 < 0 || x + 2.0 > nx || y + 2.0 > ny) {
            atomicOr(&st->bits, ERR_EDGE_REMESH);
            atomicMin(&st->first_bad, (int)k);
        }
        const uint32_t p = (nc - L) + 2 * r;
        t.ca2[p] = a; t.cb2[p] = (int32_t)k; t.cdom2[p] = t.cdom[c]; t.cbc2[p] = t.cbc[c];
        t.ca2[p + 1] = (int32_t)k; t.cb2[p + 1] = b; t.cdom2[p + 1] = t.cdom[c]; t.cbc2[p + 1] = t.cbc[c];
    }
}

// Copy the round's connections back, derive each marker's incoming/outgoing
// connection from them, give the new markers empty force rows and advance the
// counts (the workgroups read the counts the round started from).
__global__ void k_rm_commit(TopoArrays t, int32_t* min_c, int32_t* mout_c, DevStatus* st) {
    if (topo_halted(st)) return;
    const uint32_t L = st->rm_L;
    if (L == 0) return;
    const uint32_t nn = st->rm_nc0 + L, nm = st->rm_nm0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nn; j += gridDim.x * blockDim.x) {
        const int32_t a = t.ca2[j], b = t.cb2[j];
        t.ca[j] = a; t.cb[j] = b; t.cdom[j] = t.cdom2[j]; t.cbc[j] = t.cbc2[j];
        mout_c[a] = (int32_t)j;
        min_c[b] = (int32_t)j;
    }
    // new markers start with no force structs
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < L; r += gridDim.x * blockDim.x)
        t.fs_off[nm + 1 + r] = t.fs_off[nm];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rm_total += L;
        st->topo_changed = 1u;
        st->nm = nm + L;
        st->nc = nn;
        st->maxid += L;
    }
}

// ------------------------------------------------------------- rings -------
// gate != nullptr: rebuild only when *gate (the step changed the topology)
__global__ void k_ring_reset(int ndom, int32_t* start, uint32_t* cnt, const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndom) return;
    start[d] = INT_MAX;
    cnt[d] = 0;
}
// Connections of one ring mostly sit in contiguous runs of the vector: the
// head lane of each run inside a wavefront does one atomic for the whole run.
__global__ void k_ring_count(const uint32_t* __restrict__ cdom, int32_t* start, uint32_t* cnt, uint32_t* slot,
                             int ndom, DevStatus* st, const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    const uint32_t nc = st->nc;
    const uint32_t stride = gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (uint32_t c0 = blockIdx.x * blockDim.x; c0 < nc; c0 += stride) {
        const uint32_t c = c0 + threadIdx.x;
        const bool live = c < nc;
        uint32_t d = live ? cdom[c] : 0xffffffffu;
        if (live && (int)d >= ndom) { atomicOr(&st->bits, LBIB_ERR_TOPOLOGY); d = 0xfffffffeu; }
        const uint32_t dprev = __shfl_up(d, 1, 64);
        const bool head = (lane == 0) || (dprev != d);
        const unsigned long long heads = __ballot(head);
        const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
        const int hl = 63 - __clzll((long long)(heads & le));           // my run's head lane
        const unsigned long long gt = heads & ~le;                        // heads after me
        const int nextl = gt ? __ffsll((long long)gt) - 1 : 64;
        const uint32_t len = (uint32_t)(nextl - hl);
        uint32_t base = 0;
        if (head && live && d < (uint32_t)ndom) {
            base = atomicAdd(&cnt[d], len);
            atomicMin(&start[d], (int32_t)c);
        }
        base = __shfl(base, hl, 64);
        if (live && d < (uint32_t)ndom) slot[c] = base + (uint32_t)(lane - hl);
    }
}
__global__ void k_ring_bucket(const uint32_t* __restrict__ cdom, const uint32_t* __restrict__ slot,
                              const uint32_t* __restrict__ roff, int32_t* bucket, DevStatus* st, const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    const uint32_t nc = st->nc;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
        bucket[roff[cdom[c]] + slot[c]] = (int32_t)c;
}

constexpr int RING_LDS = 4096;

// one workgroup per domain: pointer jumping on the ring broken at its start
__global__ __launch_bounds__(1024) void k_ring_order(const int32_t* __restrict__ start, const uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ roff, const int32_t* __restrict__ bucket,
                                                     const uint32_t* __restrict__ slot, const int32_t* __restrict__ cb,
                                                     const int32_t* __restrict__ mout_c, int32_t* ring_marker,
                                                     DevStatus* st, const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    const int d = blockIdx.x;
    const uint32_t n = cnt[d];
    if (n == 0) return;
    if (n > (uint32_t)RING_LDS) {
        if (threadIdx.x == 0) atomicOr(&st->bits, LBIB_ERR_TOPOLOGY);
        return;
    }
    __shared__ int32_t nxt[2][RING_LDS];
    __shared__ int32_t dist[2][RING_LDS];
    __shared__ int any_live;
    const uint32_t base = roff[d];
    const int32_t s0 = start[d];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int32_t c = bucket[base + i];
        const int32_t nc_ = mout_c[cb[c]];
        nxt[0][i] = (nc_ == s0) ? -1 : (int32_t)slot[nc_];
        dist[0][i] = (nc_ == s0) ? 0 : 1;
    }
    __syncthreads();
    int cur = 0;
    for (int round = 0; round < 14; round++) {
        if (threadIdx.x == 0) any_live = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const int32_t j = nxt[cur][i];
            if (j >= 0) {
                dist[1 - cur][i] = dist[cur][i] + dist[cur][j];
                nxt[1 - cur][i] = nxt[cur][j];
                if (nxt[cur][j] >= 0) any_live = 1;
            } else {
                dist[1 - cur][i] = dist[cur][i];
                nxt[1 - cur][i] = -1;
            }
        }
        __syncthreads();
        cur = 1 - cur;
        const int live = any_live;
        __syncthreads();   // every lane has read the flag before lane 0 clears it
        if (!live) break;
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        if (nxt[cur][i] >= 0 || dist[cur][i] >= (int32_t)n) {
            atomicOr(&st->bits, LBIB_ERR_TOPOLOGY);   // open or multiple rings in one domain
            continue;
        }
        const uint32_t pos = n - 1 - (uint32_t)dist[cur][i];
        ring_marker[base + pos] = cb[bucket[base + i]];
    }
}

// Shoelace area in ring order (FP64, sequential sum as the reference's
// boost::geometry::area call site walks the ring), vertices staged in LDS.
// With mass: tutorial_01_BioSolverAreaRegulator.cpp:66-99 for the domain as
// well: default source (maxsource for cell type 1, else 0; none in domain 0),
// then for a present ring the clamped proportional source, the target being
// fixed on first use.
__device__ __forceinline__ void area_regulator_mass(const AreaMass& am, int d, bool present, double A) {
    const uint32_t ct = am.celltype[d];
    am.mtab[d] = (d == 0) ? 0.0 : ((ct == 1u) ? am.maxs : 0.0);
    if (!present) return;
    if (isnan(am.target[d])) am.target[d] = (am.factor > 0.0) ? am.factor * A : am.target_area;
    double src = am.kp * (am.target[d] - A);
    if (src > am.maxs) src = am.maxs;
    else if (src < -am.maxs) src = -am.maxs;
    if (d != 0 && ct != 1u) am.mtab[d] = src;
}

__global__ __launch_bounds__(256) void k_ring_area(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ roff,
                                                   const int32_t* __restrict__ ring_marker, const double* __restrict__ mx,
                                                   const double* __restrict__ my, double* area,
                                                   const DevStatus* gate, const AreaMass am, int with_mass) {
    if (gate && step_stopped(gate)) return;
    const int d = blockIdx.x;
    const uint32_t n = cnt[d];
    if (n == 0) {
        if (with_mass && threadIdx.x == 0) area_regulator_mass(am, d, false, 0.0);
        return;
    }
    // the shoelace terms in parallel, their sum in ring order by one lane;
    // that lane reads the terms eight at a time (four 16-byte LDS reads in
    // flight) so the chain of additions does not wait on each read
    constexpr int CH = 2048;
    __shared__ double2 term2[CH / 2];
    double* term = reinterpret_cast<double*>(term2);
    const uint32_t base = roff[d];
    double sum = 0.0;
    for (uint32_t c0 = 0; c0 < n; c0 += CH) {
        const uint32_t m = min((uint32_t)CH, n - c0);
        for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {
            const uint32_t j = c0 + k, jn = (j + 1 == n) ? 0u : j + 1;
            const int32_t q = ring_marker[base + j], r = ring_marker[base + jn];
            const double xk = mx[q], yk = my[q], xn = mx[r], yn = my[r];
            term[k] = xk * yn - xn * yk;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t k = 0;
            for (; k + 8 <= m; k += 8) {
                const double2 t0 = term2[k / 2], t1 = term2[k / 2 + 1], t2 = term2[k / 2 + 2], t3 = term2[k / 2 + 3];
                sum += t0.x; sum += t0.y; sum += t1.x; sum += t1.y;
                sum += t2.x; sum += t2.y; sum += t3.x; sum += t3.y;
            }
            for (; k < m; k++) sum += term[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        area[d] = sum * 0.5;
        if (with_mass) area_regulator_mass(am, d, true, sum * 0.5);
    }
}

// ------------------------------------------------------------- coarsen -----
__global__ void k_co_reset(DevStatus* st) {
    if (topo_halted(st)) return;
    st->co_any = 0;
    st->co_removed = 0;
}

__global__ void k_co_flag(const int32_t* __restrict__ ca, const int32_t* __restrict__ cb, const int32_t* __restrict__ min_c,
                          const int32_t* __restrict__ mout_c, const double* __restrict__ mx,
                          const double* __restrict__ my, DevStatus* st) {
    // coarsening runs only on a fully remeshed topology
    if (topo_halted(st) || (st->bits & LBIB_FLAG_REMESH)) return;
    const uint32_t nm = st->nm;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        const int p = ca[min_c[k]], q = cb[mout_c[k]];
        const double len = sqrt((mx[q] - mx[p]) * (mx[q] - mx[p]) + (my[q] - my[p]) * (my[q] - my[p]));
        if (len < kMAXLENGTH) st->co_any = 1u;
    }
}

struct CoarsenArgs {
    const uint32_t* cnt;
    const uint32_t* roff;
    const int32_t* ring_marker;   // ring members (any order is fine: sorted below)
    const int32_t *ca, *cb, *cbc, *min_c, *mout_c;
    const double *mx, *my;
    uint32_t* cdead;     // [ccap]
    uint32_t* mdead;     // [mcap]
    int32_t *rec_a, *rec_b, *rec_bc;   // [mcap] indexed by the removed marker
    uint32_t* rec_alive;               // [mcap]
};

// One workgroup per domain.  GeometryHandler::coarsenBoundary
// (GeometryHandler.cpp:433-461) visits the ring's markers in id order and
// removes a marker when its neighbours are closer than MAXLENGTH, with the
// links as left by earlier removals.  The spans of all markers are evaluated
// once in parallel with the initial links; the walk (one wavefront, in id
// order) re-evaluates only markers whose links a removal changed and
// otherwise jumps with a ballot to the next marker that may be removed.
// Positions are staged in LDS.
__global__ __launch_bounds__(256) void k_co_ring(const CoarsenArgs a, DevStatus* st) {
    if (topo_halted(st) || !st->co_any) return;
    const int d = blockIdx.x;
    const uint32_t n = a.cnt[d];
    if (n == 0) return;
    if (n > (uint32_t)RING_LDS) {
        if (threadIdx.x == 0) atomicOr(&st->bits, LBIB_ERR_TOPOLOGY);
        return;
    }
    __shared__ int32_t gid[RING_LDS];      // sorted global marker indices
    __shared__ int16_t lprev[RING_LDS], lnext[RING_LDS];
    __shared__ int32_t inc[RING_LDS], outc[RING_LDS];   // >=0 original conn, <0: -(creator local idx)-1
    __shared__ int32_t inbc[RING_LDS];
    __shared__ double sx[RING_LDS], sy[RING_LDS];
    __shared__ uint8_t flag[RING_LDS];     // bit 0: span < MAXLENGTH (initial links), bit 1: links changed
    uint32_t P = 1;
    while (P < n) P <<= 1;
    const uint32_t base = a.roff[d];
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) gid[i] = (i < n) ? a.ring_marker[base + i] : INT_MAX;
    __syncthreads();
    for (uint32_t kk = 2; kk <= P; kk <<= 1)
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
                const uint32_t l = t ^ j;
                if (l > t) {
                    const bool up = (t & kk) == 0;
                    const int32_t u = gid[t], v = gid[l];
                    if ((u > v) == up) { gid[t] = v; gid[l] = u; }
                }
            }
            __syncthreads();
        }
    auto local = [&](int32_t g) -> int {
        int lo = 0, hi = (int)n - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (gid[mid] < g) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int32_t g = gid[i];
        inc[i] = a.min_c[g];
        outc[i] = a.mout_c[g];
        inbc[i] = a.cbc[a.min_c[g]];
        lprev[i] = (int16_t)local(a.ca[a.min_c[g]]);
        lnext[i] = (int16_t)local(a.cb[a.mout_c[g]]);
        sx[i] = a.mx[g];
        sy[i] = a.my[g];
    }
    __syncthreads();
    auto span_short = [&](int p, int q) -> bool {
        const double px = sx[p], py = sy[p], qx = sx[q], qy = sy[q];
        const double len = sqrt((qx - px) * (qx - px) + (qy - py) * (qy - py));
        return len < kMAXLENGTH;
    };
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) flag[i] = span_short(lprev[i], lnext[i]) ? 1 : 0;
    __syncthreads();
    if (threadIdx.x >= 64) return;
    // one wavefront; every lane takes the same decisions and performs the
    // same LDS updates, lane 0 the global writes
    const int lane = threadIdx.x;
    uint32_t removed = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
        int from = 0;
        while (true) {
            const uint32_t i = c0 + lane;
            const bool want = i < n && lane >= from && flag[i] != 0;
            const unsigned long long m = __ballot(want);
            if (!m) break;
            const int l = __ffsll((long long)m) - 1;
            from = l + 1;
            const uint32_t ii = c0 + (uint32_t)l;
            const int p = lprev[ii], q = lnext[ii];
            const uint8_t f = flag[ii];
            const bool rm = (f & 2) ? span_short(p, q) : (f & 1) != 0;
            if (!rm) continue;
            // Geometry::removeGeometryNode (Geometry.cpp:401-428)
            const int32_t g = gid[ii];
            if (lane == 0) {
                for (int s2 = 0; s2 < 2; s2++) {
                    const int32_t cc = s2 == 0 ? inc[ii] : outc[ii];
                    if (cc >= 0) a.cdead[cc] = 1u;
                    else a.rec_alive[gid[-cc - 1]] = 0u;
                }
                a.mdead[g] = 1u;
                a.rec_a[g] = gid[p];
                a.rec_b[g] = gid[q];
                a.rec_bc[g] = inbc[ii];
                a.rec_alive[g] = 1u;
            }
            const int32_t bcv = inbc[ii];
            lnext[p] = (int16_t)q;
            lprev[q] = (int16_t)p;
            outc[p] = -(int32_t)ii - 1;
            inc[q] = -(int32_t)ii - 1;
            inbc[q] = bcv;
            flag[p] |= 2;
            flag[q] |= 2;
            removed++;
        }
    }
    if (lane == 0 && removed) {
        atomicAdd(&st->co_removed, removed);
        atomicOr(&st->topo_changed, 1u);
    }
}

// keep flags for the compaction scans
__global__ void k_co_keep(const uint32_t* __restrict__ cdead, const uint32_t* __restrict__ mdead,
                          const uint32_t* __restrict__ rec_alive, uint32_t* keepc, uint32_t* keepm, uint32_t* keepr,
                          DevStatus* st) {
    if (topo_halted(st) || !st->co_removed) return;
    const uint32_t nc = st->nc, nm = st->nm;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x)
        keepc[j] = cdead[j] ? 0u : 1u;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        keepm[k] = mdead[k] ? 0u : 1u;
        keepr[k] = (mdead[k] && rec_alive[k]) ? 1u : 0u;
    }
}

// build the compacted connection list into the scratch arrays:
// surviving originals in order, then alive records in removed-marker order
__global__ void k_co_conns(TopoArrays t, const uint32_t* __restrict__ keepc, const uint32_t* __restrict__ posc,
                           const uint32_t* __restrict__ keepr, const uint32_t* __restrict__ posr,
                           const uint32_t* __restrict__ newm, const int32_t* __restrict__ rec_a,
                           const int32_t* __restrict__ rec_b, const int32_t* __restrict__ rec_bc,
                           const uint32_t* __restrict__ mdom, const uint32_t* nkeepc, DevStatus* st) {
    if (topo_halted(st) || !st->co_removed) return;
    const uint32_t nc = st->nc, nm = st->nm, ns = *nkeepc;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x) {
        if (!keepc[j]) continue;
        const uint32_t p = posc[j];
        t.ca2[p] = (int32_t)newm[t.ca[j]];
        t.cb2[p] = (int32_t)newm[t.cb[j]];
        t.cdom2[p] = t.cdom[j];
        t.cbc2[p] = t.cbc[j];
    }
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        if (!keepr[k]) continue;
        const uint32_t p = ns + posr[k];
        t.ca2[p] = (int32_t)newm[rec_a[k]];
        t.cb2[p] = (int32_t)newm[rec_b[k]];
        t.cdom2[p] = mdom[k];
        t.cbc2[p] = rec_bc[k];
    }
}

// domain of each marker = domain of its incoming connection (before compaction)
__global__ void k_marker_dom(const int32_t* __restrict__ min_c, const uint32_t* __restrict__ cdom, uint32_t* mdom,
                             DevStatus* st) {
    if (topo_halted(st) || !st->co_removed) return;
    const uint32_t nm = st->nm;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) mdom[k] = cdom[min_c[k]];
}

struct MarkerScratch {
    double *x, *y, *vx, *vy, *ox, *oy;
    uint32_t* id;
};

__global__ void k_co_markers(TopoArrays t, MarkerScratch s, const uint32_t* __restrict__ keepm,
                             const uint32_t* __restrict__ newm, DevStatus* st) {
    if (topo_halted(st) || !st->co_removed) return;
    const uint32_t nm = st->nm;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        if (!keepm[k]) continue;
        const uint32_t p = newm[k];
        s.x[p] = t.mx[k]; s.y[p] = t.my[k]; s.vx[p] = t.mvx[k]; s.vy[p] = t.mvy[k];
        s.ox[p] = t.mox[k]; s.oy[p] = t.moy[k]; s.id[p] = t.mid[k];
    }
}

__global__ void k_co_commit(TopoArrays t, MarkerScratch s, const uint32_t* nkeepm, const uint32_t* nkeepc,
                            const uint32_t* nkeepr, DevStatus* st) {
    if (topo_halted(st) || !st->co_removed) return;
    const uint32_t nm = *nkeepm, nc = *nkeepc + *nkeepr;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        t.mx[k] = s.x[k]; t.my[k] = s.y[k]; t.mvx[k] = s.vx[k]; t.mvy[k] = s.vy[k];
        t.mox[k] = s.ox[k]; t.moy[k] = s.oy[k]; t.mid[k] = s.id[k];
    }
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x) {
        t.ca[j] = t.ca2[j]; t.cb[j] = t.cb2[j]; t.cdom[j] = t.cdom2[j]; t.cbc[j] = t.cbc2[j];
    }
}

__global__ void k_co_counts(const uint32_t* nkeepm, const uint32_t* nkeepc, const uint32_t* nkeepr, DevStatus* st) {
    if (topo_halted(st) || !st->co_removed) return;
    st->nm = *nkeepm;
    st->nc = *nkeepc + *nkeepr;
}

__global__ void k_clear_u32(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// ForceSolver::deleteAllForces (ForceSolver.cpp:158-166): empty CSR rows
__global__ void k_forces_clear(int32_t* fs_off, DevStatus* st) {
    if (step_stopped(st)) return;
    const uint32_t nm = st->nm;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= nm; k += gridDim.x * blockDim.x) fs_off[k] = 0;
}

// tutorial_01_BioSolverMembraneTension (…MembraneTension.cpp:53-76; type 6,
// p0 = f) and BioSolverHookeanMembraneTension (…HookeanMembraneTension.cpp:
// 55-80; type 0, p0 = K, p1 = l_resting) over a table without other structs:
// every marker holds exactly its two structs, one per adjacent connection;
// with two terms the summation order does not matter.
__global__ void k_forces_pairs(const int32_t* __restrict__ ca, const int32_t* __restrict__ cb,
                               const int32_t* __restrict__ min_c, const int32_t* __restrict__ mout_c, int type,
                               double p0, double p1, int32_t* fs_off, int32_t* fs_type, int32_t* fs_partner,
                               double* fs_prm, DevStatus* st) {
    if (step_stopped(st)) return;
    const uint32_t nm = st->nm;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        fs_off[k] = (int32_t)(2 * k);
        if (k + 1 == nm) fs_off[nm] = (int32_t)(2 * nm);
        const uint32_t e = 2 * k;
        fs_type[e] = type; fs_partner[e] = cb[mout_c[k]];
        fs_type[e + 1] = type; fs_partner[e + 1] = ca[min_c[k]];
        for (int q = 0; q < 4; q++) { fs_prm[4 * e + q] = 0.0; fs_prm[4 * e + 4 + q] = 0.0; }
        fs_prm[4 * e] = p0;
        fs_prm[4 * e + 1] = p1;
        fs_prm[4 * e + 4] = p0;
        fs_prm[4 * e + 5] = p1;
    }
}

// tutorial_02_BioSolverGrowth (tutorial_02_BioSolverGrowth.cpp:54-80): on
// every cell site addMass(R*R*L > threshold ? maxsource : maxsource/10) with
// R, L the getC() of the two species (summed in the locally swapped slot
// order); other sites keep the mass an earlier BioSolver of this iteration
// gave them (addMass overwrites).
template <int P>
__global__ void k_mass_growth(long n, const void* g, int sR, int sL, int nq, const uint32_t* __restrict__ tag,
                              double threshold, double maxsource, int keep_site, const double* __restrict__ mtab,
                              int ntab, double* msite, const DevStatus* st) {
    if (step_stopped(st)) return;
    using GS = GStore<P>;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const uint32_t d = tag[i] & DOM_MASK;
        if (d == 0) {
            if (!keep_site) msite[i] = (mtab && (int)d < ntab) ? mtab[d] : 0.0;
            continue;
        }
        auto conc = [&](int s) {
            return c_swapped(GS::get(g, s, 0, i, n, nq), GS::get(g, s, 1, i, n, nq), GS::get(g, s, 2, i, n, nq),
                             GS::get(g, s, 3, i, n, nq), GS::get(g, s, 4, i, n, nq));
        };
        const double R = conc(sR), L = conc(sL);
        msite[i] = (R * R * L > threshold) ? maxsource : maxsource / 10.0;
    }
}

// FluidSolver::addMass at a list of sites (FluidSolver.cpp:253-257: the call
// overwrites, so the last value given for a site this iteration counts).  The
// pending mass of the iteration is first made per-site (k_mass_fill: the
// per-domain table, or zero, where no per-site field is pending yet), then the
// listed sites are overwritten in list order by one thread (k_mass_scatter's
// grid-stride loop keeps the last write of a repeated site when sites are
// unique per list; duplicates are resolved on the host).
__global__ void k_mass_fill(long n, const uint32_t* __restrict__ tag, const double* __restrict__ mtab, int ntab,
                            double* msite, const DevStatus* st) {
    if (step_stopped(st)) return;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const uint32_t d = tag[i] & DOM_MASK;
        msite[i] = (mtab && (int)d < ntab) ? mtab[d] : 0.0;
    }
}
__global__ void k_mass_scatter(long cnt, const int64_t* __restrict__ site, const double* __restrict__ val, double* msite,
                               const DevStatus* st) {
    if (step_stopped(st)) return;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += (long)gridDim.x * blockDim.x)
        msite[site[k]] = val[k];
}

// --------------------------------------------------------------- drivers ---
static inline int grid_for(long n) { return (int)std::min<long>(std::max<long>((n + 255) / 256, 1), 4096); }

static TopoArrays topo_arrays(lbib_ctx* c) {
    TopoArrays t;
    t.mx = c->mx; t.my = c->my; t.mvx = c->mvx; t.mvy = c->mvy; t.mox = c->mox; t.moy = c->moy; t.mid = c->mid;
    t.ca = c->ca; t.cb = c->cb; t.cbc = c->cbc; t.cdom = c->cdom;
    t.ca2 = c->ca2; t.cb2 = c->cb2; t.cbc2 = c->cbc2; t.cdom2 = c->cdom2;
    t.fs_off = c->fs_off;
    return t;
}

int launch_inout(lbib_ctx* c) {
    k_inout<<<grid_for(c->ccap), 256, 0, c->stream>>>(c->ca, c->cb, c->min_c, c->mout_c, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_rings(lbib_ctx* c, bool counters_cleared, bool gated) {
    if (c->ndom == 0) return 0;
    const uint32_t* gate = gated ? &c->st->topo_changed : nullptr;
    if (!counters_cleared)
        k_ring_reset<<<(c->ndom + 255) / 256, 256, 0, c->stream>>>(c->ndom, c->ring_start, c->ring_cnt, gate);
    k_ring_count<<<grid_for(c->ccap), 256, 0, c->stream>>>(c->cdom, c->ring_start, c->ring_cnt, c->ring_slot, c->ndom,
                                                          c->st, gate);
    if (scan_exclusive(c, c->ring_cnt, c->ring_off_u, nullptr, (uint32_t)c->ndom, (uint32_t)c->ndom, nullptr, gate))
        return -1;
    k_ring_bucket<<<grid_for(c->ccap), 256, 0, c->stream>>>(c->cdom, c->ring_slot, c->ring_off_u, c->ring_bucket,
                                                           c->st, gate);
    k_ring_order<<<c->ndom, 1024, 0, c->stream>>>(c->ring_start, c->ring_cnt, c->ring_off_u, c->ring_bucket,
                                                  c->ring_slot, c->cb, c->mout_c, c->ring_marker, c->st, gate);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_ring_areas(lbib_ctx* c, const AreaMass* mass) {
    if (c->ndom == 0) return 0;
    const AreaMass none{};
    k_ring_area<<<c->ndom, 256, 0, c->stream>>>(c->ring_cnt, c->ring_off_u, c->ring_marker, c->mx, c->my, c->area,
                                                mass ? c->st : nullptr, mass ? *mass : none, mass ? 1 : 0);
    LBIB_HIP(hipGetLastError());
    return 0;
}

// Rounds of remeshing; LBIB_FLAG_REMESH asks for another pass when a
// connection is still long (a marker moves far less than MAXLENGTH per
// iteration, so one round is the steady-state case).
int launch_remesh(lbib_ctx* c, int rounds) {
    TopoArrays t = topo_arrays(c);
    for (int r = 0; r < rounds; r++) {
        k_rm_flag<<<grid_for(c->ccap), 256, 0, c->stream>>>(c->ca, c->cb, c->mx, c->my, c->flag_c, c->rm_list, 0,
                                                            nullptr, nullptr, c->st);
        k_rm_rank<<<1, 1024, 0, c->stream>>>(c->flag_c, c->rank_c, c->rm_list, c->st);
        k_rm_apply<<<grid_for(c->ccap), 256, 0, c->stream>>>(t, c->flag_c, c->rank_c, c->rm_list, c->nx, c->ny,
                                                             (uint32_t)c->mcap, (uint32_t)c->ccap, c->st);
        k_rm_commit<<<grid_for(2 * (long)c->ccap), 256, 0, c->stream>>>(t, c->min_c, c->mout_c, c->st);
    }
    // with the clearing of the ring counters for the rebuild that follows
    k_rm_flag<<<grid_for(c->ccap), 256, 0, c->stream>>>(c->ca, c->cb, c->mx, c->my, nullptr, nullptr, c->ndom,
                                                        c->ring_start, c->ring_cnt, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_coarsen(lbib_ctx* c) {
    TopoArrays t = topo_arrays(c);
    k_co_reset<<<1, 1, 0, c->stream>>>(c->st);
    k_co_flag<<<grid_for(c->mcap), 256, 0, c->stream>>>(c->ca, c->cb, c->min_c, c->mout_c, c->mx, c->my, c->st);
    k_clear_u32<<<grid_for(c->ccap), 256, 0, c->stream>>>(c->cdead, (uint32_t)c->ccap);
    k_clear_u32<<<grid_for(c->mcap), 256, 0, c->stream>>>(c->mdead, (uint32_t)c->mcap);
    CoarsenArgs a;
    a.cnt = c->ring_cnt; a.roff = c->ring_off_u; a.ring_marker = c->ring_marker;
    a.ca = c->ca; a.cb = c->cb; a.cbc = c->cbc; a.min_c = c->min_c; a.mout_c = c->mout_c;
    a.mx = c->mx; a.my = c->my; a.cdead = c->cdead; a.mdead = c->mdead;
    a.rec_a = c->rec_a; a.rec_b = c->rec_b; a.rec_bc = c->rec_bc; a.rec_alive = c->rec_alive;
    if (c->ndom > 0) k_co_ring<<<c->ndom, 256, 0, c->stream>>>(a, c->st);
    k_marker_dom<<<grid_for(c->mcap), 256, 0, c->stream>>>(c->min_c, c->cdom, c->mdom, c->st);
    k_co_keep<<<grid_for(std::max(c->mcap, c->ccap)), 256, 0, c->stream>>>(c->cdead, c->mdead, c->rec_alive, c->keep_c,
                                                                          c->keep_m, c->keep_r, c->st);
    if (scan_exclusive(c, c->keep_c, c->pos_c, &c->st->nc, 0, (uint32_t)c->ccap, &c->st->co_nkeepc)) return -1;
    if (scan_exclusive(c, c->keep_m, c->pos_m, &c->st->nm, 0, (uint32_t)c->mcap, &c->st->co_nkeepm)) return -1;
    if (scan_exclusive(c, c->keep_r, c->pos_r, &c->st->nm, 0, (uint32_t)c->mcap, &c->st->co_nkeepr)) return -1;
    k_co_conns<<<grid_for(std::max(c->mcap, c->ccap)), 256, 0, c->stream>>>(
        t, c->keep_c, c->pos_c, c->keep_r, c->pos_r, c->pos_m, c->rec_a, c->rec_b, c->rec_bc, c->mdom,
        &c->st->co_nkeepc, c->st);
    MarkerScratch s;
    s.x = c->sx_m; s.y = c->sy_m; s.vx = c->svx_m; s.vy = c->svy_m; s.ox = c->sox_m; s.oy = c->soy_m; s.id = c->sid_m;
    k_co_markers<<<grid_for(c->mcap), 256, 0, c->stream>>>(t, s, c->keep_m, c->pos_m, c->st);
    k_co_commit<<<grid_for(std::max(c->mcap, c->ccap)), 256, 0, c->stream>>>(t, s, &c->st->co_nkeepm,
                                                                            &c->st->co_nkeepc, &c->st->co_nkeepr, c->st);
    k_co_counts<<<1, 1, 0, c->stream>>>(&c->st->co_nkeepm, &c->st->co_nkeepc, &c->st->co_nkeepr, c->st);
    LBIB_HIP(hipGetLastError());
    if (launch_inout(c)) return -1;
    return 0;
}

int launch_forces_clear(lbib_ctx* c) {
    k_forces_clear<<<grid_for(c->mcap + 1), 256, 0, c->stream>>>(c->fs_off, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_forces_pairs(lbib_ctx* c, int type, double p0, double p1) {
    k_forces_pairs<<<grid_for(c->mcap), 256, 0, c->stream>>>(c->ca, c->cb, c->min_c, c->mout_c, type, p0, p1,
                                                             c->fs_off, c->fs_type, c->fs_partner, c->fs_prm, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

namespace {
__global__ void k_pert_collect(int nm, const double* __restrict__ mx, const double* __restrict__ my,
                               const uint32_t* __restrict__ mid, int cap, uint32_t* cnt, int32_t* ix, uint32_t* id,
                               double* xy) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        const double X = mx[k], Y = my[k];
        if (!needs_perturbation(X, Y)) continue;
        const uint32_t s = atomicAdd(cnt, 1u);
        if (s < (uint32_t)cap) {
            ix[s] = k;
            id[s] = mid[k];
            xy[2 * s] = X;
            xy[2 * s + 1] = Y;
        }
    }
}
__global__ void k_conn_pert_domains(int nc, int ndom, const int32_t* __restrict__ ca, const int32_t* __restrict__ cb,
                                    const uint32_t* __restrict__ cdom, const double* __restrict__ mx,
                                    const double* __restrict__ my, uint32_t* dflag, int cap, uint32_t* cnt,
                                    uint32_t* dlist) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const int a = ca[c], b = cb[c];
        if (!(fabs(mx[a] - mx[b]) < kEPS || fabs(my[a] - my[b]) < kEPS)) continue;
        const uint32_t d = cdom[c];
        if (d >= (uint32_t)ndom || atomicExch(&dflag[d], 1u) != 0u) continue;
        const uint32_t s = atomicAdd(cnt, 1u);
        if (s < (uint32_t)cap) dlist[s] = d;
    }
}
__global__ void k_ring_info(int nd, const uint32_t* __restrict__ dlist, const uint32_t* __restrict__ ring_cnt,
                            const uint32_t* __restrict__ ring_off, uint32_t* len, uint32_t* base) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += gridDim.x * blockDim.x) {
        len[i] = ring_cnt[dlist[i]];
        base[i] = ring_off[dlist[i]];
    }
}
// one workgroup per listed ring; ring_marker[base + p] is the second marker of
// the p-th connection from the ring's first, so the first marker of
// connection j is ring_marker[base + j - 1] (j = 0: the last entry)
__global__ void k_ring_gather(const uint32_t* __restrict__ base, const int32_t* __restrict__ off,
                              const int32_t* __restrict__ ring_marker, const double* __restrict__ mx,
                              const double* __restrict__ my, int32_t* marker, double* xy) {
    const int i = blockIdx.x;
    const int o = off[i], n = off[i + 1] - off[i];
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int32_t m = ring_marker[base[i] + (uint32_t)(j == 0 ? n - 1 : j - 1)];
        marker[o + j] = m;
        xy[2 * (o + j)] = mx[m];
        xy[2 * (o + j) + 1] = my[m];
    }
}
__global__ void k_set_positions(int n, const int32_t* __restrict__ ix, const double* __restrict__ xy, double* mx,
                                double* my) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        mx[ix[i]] = xy[2 * i];
        my[ix[i]] = xy[2 * i + 1];
    }
}
}  // namespace

int launch_perturbation_candidates(lbib_ctx* c, int cap, uint32_t* cnt, int32_t* ix, uint32_t* id, double* xy) {
    if (c->nm <= 0) return 0;
    k_pert_collect<<<grid_for(c->nm), 256, 0, c->stream>>>(c->nm, c->mx, c->my, c->mid, cap, cnt, ix, id, xy);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_conn_perturbation_domains(lbib_ctx* c, uint32_t* dflag, int cap, uint32_t* cnt, uint32_t* dlist) {
    if (c->nc <= 0) return 0;
    k_conn_pert_domains<<<grid_for(c->nc), 256, 0, c->stream>>>(c->nc, c->ndom, c->ca, c->cb, c->cdom, c->mx, c->my,
                                                                dflag, cap, cnt, dlist);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_ring_info(lbib_ctx* c, int nd, const uint32_t* dlist, uint32_t* len, uint32_t* base) {
    if (nd <= 0) return 0;
    k_ring_info<<<grid_for(nd), 256, 0, c->stream>>>(nd, dlist, c->ring_cnt, c->ring_off_u, len, base);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_ring_gather(lbib_ctx* c, int nd, const uint32_t* base, const int32_t* off, int32_t* marker, double* xy) {
    if (nd <= 0) return 0;
    k_ring_gather<<<nd, 256, 0, c->stream>>>(base, off, c->ring_marker, c->mx, c->my, marker, xy);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_set_positions(lbib_ctx* c, int n, const int32_t* ix, const double* xy) {
    if (n <= 0) return 0;
    k_set_positions<<<grid_for(n), 256, 0, c->stream>>>(n, ix, xy, c->mx, c->my);
    LBIB_HIP(hipGetLastError());
    return 0;
}

int launch_mass_sites(lbib_ctx* c, long cnt, const int64_t* site, const double* val) {
    if (!c->msite_live) {
        const double* mt = c->mtab_live ? c->mtab : nullptr;
        k_mass_fill<<<grid_for(c->n), 256, 0, c->stream>>>(c->n, c->tag, mt, c->ndom, c->msite, c->st);
        LBIB_HIP(hipGetLastError());
    }
    if (cnt > 0) {
        k_mass_scatter<<<grid_for(cnt), 256, 0, c->stream>>>(cnt, site, val, c->msite, c->st);
        LBIB_HIP(hipGetLastError());
    }
    c->msite_live = true;
    return 0;
}

int launch_mass_growth(lbib_ctx* c, int sR, int sL, double threshold, double maxsource) {
    const void* g = c->g[c->cur];
    const int nq = 5 * std::max(c->ns, 1);
    const int keep = c->msite_live ? 1 : 0;
    const double* mt = c->mtab_live ? c->mtab : nullptr;
    const int gs = (int)std::min<long>((c->n + 255) / 256, 8192);
    switch (c->prec) {
        case LBIB_FP64:
            k_mass_growth<LBIB_FP64><<<gs, 256, 0, c->stream>>>(c->n, g, sR, sL, nq, c->tag, threshold, maxsource, keep,
                                                                mt, c->ndom, c->msite, c->st);
            break;
        case LBIB_FP32:
            k_mass_growth<LBIB_FP32><<<gs, 256, 0, c->stream>>>(c->n, g, sR, sL, nq, c->tag, threshold, maxsource, keep,
                                                                mt, c->ndom, c->msite, c->st);
            break;
        case LBIB_FP16_DEV:
            k_mass_growth<LBIB_FP16_DEV><<<gs, 256, 0, c->stream>>>(c->n, g, sR, sL, nq, c->tag, threshold, maxsource,
                                                                    keep, mt, c->ndom, c->msite, c->st);
            break;
        case LBIB_FP16_REM:
            k_mass_growth<LBIB_FP16_REM><<<gs, 256, 0, c->stream>>>(c->n, g, sR, sL, nq, c->tag, threshold, maxsource,
                                                                    keep, mt, c->ndom, c->msite, c->st);
            break;
        default:
            k_mass_growth<LBIB_BF16_DEV><<<gs, 256, 0, c->stream>>>(c->n, g, sR, sL, nq, c->tag, threshold, maxsource,
                                                                    keep, mt, c->ndom, c->msite, c->st);
            break;
    }
    LBIB_HIP(hipGetLastError());
    c->msite_live = true;
    return 0;
}

// GeometryHandler::computeAccumulatedDomainConcentrations
// (GeometryHandler.cpp:136-147): the sum of getC() of one species over the
// sites of each domain, x from 1 (the reference skips the x = 0 column).
// One workgroup per row; runs of equal domain inside a wavefront are summed
// with a segmented scan and added once per run (domain 0 in LDS, one global
// atomic per row).
template <int P>
__global__ __launch_bounds__(256) void k_domain_conc(int nx, long n, const void* g, int species, int nq,
                                                    const uint32_t* __restrict__ tag, int ndom, double* acc) {
    using GS = GStore<P>;
    __shared__ double out0;
    if (threadIdx.x == 0) out0 = 0.0;
    __syncthreads();
    const long y = blockIdx.x;
    const int lane = threadIdx.x & 63;
    for (int x0 = 1; x0 < nx; x0 += blockDim.x) {
        const int x = x0 + (int)threadIdx.x;
        uint32_t d = 0xffffffffu;
        double s = 0.0;
        if (x < nx) {
            const long i = y * nx + x;
            d = tag[i] & DOM_MASK;
            s = c_swapped(GS::get(g, species, 0, i, n, nq), GS::get(g, species, 1, i, n, nq),
                          GS::get(g, species, 2, i, n, nq), GS::get(g, species, 3, i, n, nq),
                          GS::get(g, species, 4, i, n, nq));
        }
        const uint32_t dprev = (uint32_t)__shfl_up((int)d, 1, 64);
        int head = (lane == 0 || dprev != d) ? 1 : 0;
        for (int o = 1; o < 64; o <<= 1) {
            const double ts = __shfl_up(s, o, 64);
            const int th = __shfl_up(head, o, 64);
            if (lane >= o) {
                if (!head) s += ts;
                head |= th;
            }
        }
        const uint32_t dnext = (uint32_t)__shfl_down((int)d, 1, 64);
        const bool tail = lane == 63 || dnext != d;
        if (tail && d != 0xffffffffu) {
            if (d == 0u) atomicAdd(&out0, s);
            else if ((int)d < ndom) atomicAdd(&acc[d], s);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&acc[0], out0);
}

int launch_domain_concentrations(lbib_ctx* c, int species, int ndom, double* acc) {
    const void* g = c->g[c->cur];
    const int nq = 5 * std::max(c->ns, 1);
    switch (c->prec) {
        case LBIB_FP64: k_domain_conc<LBIB_FP64><<<c->ny, 256, 0, c->stream>>>(c->nx, c->n, g, species, nq, c->tag, ndom, acc); break;
        case LBIB_FP32: k_domain_conc<LBIB_FP32><<<c->ny, 256, 0, c->stream>>>(c->nx, c->n, g, species, nq, c->tag, ndom, acc); break;
        case LBIB_FP16_DEV:
            k_domain_conc<LBIB_FP16_DEV><<<c->ny, 256, 0, c->stream>>>(c->nx, c->n, g, species, nq, c->tag, ndom, acc);
            break;
        case LBIB_FP16_REM:
            k_domain_conc<LBIB_FP16_REM><<<c->ny, 256, 0, c->stream>>>(c->nx, c->n, g, species, nq, c->tag, ndom, acc);
            break;
        default: k_domain_conc<LBIB_BF16_DEV><<<c->ny, 256, 0, c->stream>>>(c->nx, c->n, g, species, nq, c->tag, ndom, acc); break;
    }
    LBIB_HIP(hipGetLastError());
    return 0;
}

}  // namespace lbib
