// Cell junctions on the device force table: tutorial_02_BioSolverCellJunction
// (tutorial_02_BioSolverCellJunction.cpp:47-99, BioSolverCellJunction.cpp is
// the same process).  Every application
//   1. deletes the type-7 structs (ForceSolver::deleteForceType,
//      ForceSolver.cpp:168-204; the other structs keep their order),
//   2. for each marker j in GeometryNode id order finds the closest marker c
//      within RADIUS whose adjacent connections belong to another domain
//      (Geometry::getGeometryNodesWithinRadiusWithAvoidanceClosest,
//      FastNeighborList.hpp:404-467: distance sqrt(dx*dx + dy*dy) <= RADIUS),
//   3. appends the spring (7, j, c, K, L0) to j's row and (7, c, j, K, L0)
//      to c's row (ForceSolver::addForce pushes to the back of a row).
// A marker's row therefore ends with its junctions ordered by the id of the
// marker whose search created them: its own search (partner c) and every
// marker s that chose it (partner s).  Equidistant candidates resolve to the
// smaller id; the reference takes the first in its bin traversal order, the
// same marker whenever both share a neighbour-list bin.
//
// Device plan: markers are binned by their 16x16 tile; one thread per marker
// scans the tiles its search box touches, records its partner and registers
// itself with the partner (at most JUNCTION_MAXIN finders per marker, beyond
// that LBIB_ERR_CAPACITY).  New row lengths are scanned into offsets, each
// marker assembles its row in a scratch CSR (kept structs, then junctions by
// finder index) and the scratch is copied over the table.  Marker index order
// is GeometryNode id order.
#include "lbib_internal.h"

namespace lbib {
namespace {

constexpr int JT = 16;   // tile edge of the marker bins (the spread tiles)

__global__ void k_junction_clear(uint32_t* in, int n, const DevStatus* st) {
    if (step_stopped(st)) return;
    const uint32_t nm = st->nm;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nm && k < (uint32_t)n;
         k += gridDim.x * blockDim.x)
        in[k] = 0u;
}

// GeometryNode::getDomainIdOfAdjacentConnections (GeometryNode.cpp:267-283):
// the incoming connection's domain (both adjacent connections carry it)
__global__ void k_junction_find(int ntx, int nty, double radius, const double* __restrict__ mx,
                                const double* __restrict__ my, const uint32_t* __restrict__ tcnt,
                                const int32_t* __restrict__ tlist, const BinOvf* __restrict__ ovf,
                                const uint32_t* __restrict__ cdom, const int32_t* __restrict__ min_c,
                                const int32_t* __restrict__ rtile, int32_t* __restrict__ part, uint32_t* in_cnt,
                                int32_t* __restrict__ in_src, DevStatus* st) {
    if (step_stopped(st)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int)st->nm) return;
    const double x = mx[k], y = my[k];
    const uint32_t avoid = cdom[min_c[k]];
    const int fxl = (int)floor(x - radius), fxh = (int)floor(x + radius);
    const int fyl = (int)floor(y - radius), fyh = (int)floor(y + radius);
    const int tx0 = fxl < 0 ? 0 : fxl / JT, ty0 = fyl < 0 ? 0 : fyl / JT;
    const int tx1 = min(fxh < 0 ? -1 : fxh / JT, ntx - 1), ty1 = min(fyh < 0 ? -1 : fyh / JT, nty - 1);
    // independent domains of a batched atlas never pair across rectangles
    const int hx = min(max((int)floor(x), 0) / JT, ntx - 1), hy = min(max((int)floor(y), 0) / JT, nty - 1);
    const int rq = rtile ? rtile[hy * ntx + hx] : -1;
    int best = -1;
    double bd = 0.0;
    auto consider = [&](int q) {
        const double dx = mx[q] - x, dy = my[q] - y;
        const double d = sqrt(dx * dx + dy * dy);
        if (d > radius) return;
        if (cdom[min_c[q]] == avoid) return;
        if (best < 0 || d < bd || (d == bd && q < best)) {
            best = q;
            bd = d;
        }
    };
    bool over = false;
    for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1; tx++) {
            const int t = ty * ntx + tx;
            if (rtile && rtile[t] != rq) continue;
            const uint32_t n = tcnt[t];
            over |= n > (uint32_t)BINCAP;
            const long b = (long)t * BINCAP;
            for (uint32_t i = 0; i < min(n, (uint32_t)BINCAP); i++) consider(tlist[b + i]);
        }
    // entries past a tile's capacity
    if (over) {
        const uint32_t no = min(st->novf, OVF_CAP);
        for (uint32_t j = 0; j < no; j++) {
            const int t = ovf[j].tile, tx = t % ntx, ty = t / ntx;
            if (tx < tx0 || tx > tx1 || ty < ty0 || ty > ty1) continue;
            if (rtile && rtile[t] != rq) continue;
            consider(ovf[j].k);
        }
    }
    part[k] = best;
    if (best >= 0) {
        const uint32_t slot = atomicAdd(&in_cnt[best], 1u);
        if (slot < (uint32_t)JUNCTION_MAXIN) in_src[(long)best * JUNCTION_MAXIN + slot] = k;
        else atomicOr(&st->bits, LBIB_ERR_CAPACITY);
    }
}

// new row length: the row's structs of other types plus its junctions
__global__ void k_junction_count(const int32_t* __restrict__ off, const int32_t* __restrict__ type,
                                 const int32_t* __restrict__ part, const uint32_t* __restrict__ in_cnt,
                                 uint32_t* __restrict__ cnt, const DevStatus* st) {
    if (step_stopped(st)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int)st->nm) return;
    uint32_t n = 0;
    for (int j = off[k]; j < off[k + 1]; j++) n += type[j] != 7;
    n += (part[k] >= 0 ? 1u : 0u) + min(in_cnt[k], (uint32_t)JUNCTION_MAXIN);
    cnt[k] = n;
}

__global__ void k_junction_fill(const int32_t* __restrict__ off, const int32_t* __restrict__ type,
                                const int32_t* __restrict__ partner, const double* __restrict__ prm,
                                const int32_t* __restrict__ part, const uint32_t* __restrict__ in_cnt,
                                const int32_t* __restrict__ in_src, const uint32_t* __restrict__ ofs,
                                const uint32_t* __restrict__ total, int fcap, double K, double L0,
                                int32_t* __restrict__ noff, int32_t* __restrict__ ntype,
                                int32_t* __restrict__ npartner, double* __restrict__ nprm, DevStatus* st) {
    if (step_stopped(st)) return;
    const uint32_t nf = *total;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (nf > (uint32_t)fcap) {
        if (k == 0) atomicOr(&st->bits, LBIB_ERR_CAPACITY);
        return;
    }
    const int nm = (int)st->nm;
    if (k >= nm) return;
    uint32_t o = ofs[k];
    noff[k] = (int32_t)o;
    if (k == nm - 1) noff[nm] = (int32_t)nf;
    for (int j = off[k]; j < off[k + 1]; j++) {
        if (type[j] == 7) continue;
        ntype[o] = type[j];
        npartner[o] = partner[j];
        for (int q = 0; q < 4; q++) nprm[4 * (long)o + q] = prm[4 * (long)j + q];
        o++;
    }
    // finders in index order: this marker's own search sits among them at k
    int src[JUNCTION_MAXIN + 1];
    int ns = 0;
    const int nin = (int)min(in_cnt[k], (uint32_t)JUNCTION_MAXIN);
    for (int i = 0; i < nin; i++) {
        const int s = in_src[(long)k * JUNCTION_MAXIN + i];
        int p = ns++;
        while (p > 0 && src[p - 1] > s) {
            src[p] = src[p - 1];
            p--;
        }
        src[p] = s;
    }
    const int own = part[k];
    bool placed = own < 0;
    int i = 0;
    while (i < ns || !placed) {
        int q;
        if (!placed && (i == ns || src[i] > k)) {
            q = own;
            placed = true;
        } else {
            q = src[i++];
        }
        ntype[o] = 7;
        npartner[o] = q;
        nprm[4 * (long)o] = K;
        nprm[4 * (long)o + 1] = L0;
        nprm[4 * (long)o + 2] = 0.0;
        nprm[4 * (long)o + 3] = 0.0;
        o++;
    }
}

__global__ void k_junction_commit(const int32_t* __restrict__ noff, const int32_t* __restrict__ ntype,
                                  const int32_t* __restrict__ npartner, const double* __restrict__ nprm,
                                  const uint32_t* __restrict__ total, int fcap, int32_t* __restrict__ off,
                                  int32_t* __restrict__ type, int32_t* __restrict__ partner,
                                  double* __restrict__ prm, const DevStatus* st) {
    if (step_stopped(st)) return;
    const uint32_t nf = *total;
    if (nf > (uint32_t)fcap) return;
    const uint32_t nm = st->nm;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= nm; i += stride) off[i] = noff[i];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nf; j += stride) {
        type[j] = ntype[j];
        partner[j] = npartner[j];
        for (int q = 0; q < 4; q++) prm[4 * (long)j + q] = nprm[4 * (long)j + q];
    }
}

inline int blocks_for(long n, int bs) { return (int)std::max<long>((n + bs - 1) / bs, 1); }

}  // namespace

int launch_forces_junctions(lbib_ctx* c, double radius, double k, double l0) {
    const int bs = 256;
    if (launch_tile_bins(c)) return -1;
    k_junction_clear<<<blocks_for(c->mcap, bs), bs, 0, c->stream>>>(c->j_in, c->jcap, c->st);
    k_junction_find<<<blocks_for(c->mcap, bs), bs, 0, c->stream>>>(
        c->stx, c->sty, radius, c->mx, c->my, c->tcnt, c->tlist, c->ovf, c->cdom, c->min_c,
        c->nrect ? c->rtile : nullptr, c->j_part, c->j_in, c->j_src, c->st);
    k_junction_count<<<blocks_for(c->mcap, bs), bs, 0, c->stream>>>(c->fs_off, c->fs_type, c->j_part, c->j_in,
                                                                     c->j_cnt, c->st);
    LBIB_HIP(hipGetLastError());
    if (scan_exclusive(c, c->j_cnt, c->j_ofs, &c->st->nm, 0, (uint32_t)c->mcap, c->j_total)) return -1;
    k_junction_fill<<<blocks_for(c->mcap, bs), bs, 0, c->stream>>>(
        c->fs_off, c->fs_type, c->fs_partner, c->fs_prm, c->j_part, c->j_in, c->j_src, c->j_ofs, c->j_total,
        c->jfcap, k, l0, c->j_off, c->j_type, c->j_partner, c->j_prm, c->st);
    k_junction_commit<<<std::min(blocks_for((long)c->fcap, bs), 4096), bs, 0, c->stream>>>(
        c->j_off, c->j_type, c->j_partner, c->j_prm, c->j_total, std::min(c->fcap, c->jfcap), c->fs_off,
        c->fs_type, c->fs_partner, c->fs_prm, c->st);
    LBIB_HIP(hipGetLastError());
    return 0;
}

}  // namespace lbib
