// lbib_internal.h -- device-side data layout and shared helpers of liblbib_hip.
//
// HBM layout of one context (n = nx*ny sites, row-major, x fastest):
//   f[2][9][n]        populations, natural direction order, post-collision
//                     (A/B ping-pong; pull streaming reads cur, writes 1-cur)
//   g[2][ns][5][n]    species populations, same convention
//   u[n][2]           FluidSolver::velocity_ (post-collision), (ux, uy) per site
//                     so that one 16-byte access serves an interpolation stencil site
//   Fx[n], Fy[n]      FluidSolver::force_ (IB spread target, zero off-membrane)
//   tag[n]            bits 0-23 domain id, 24-27 boundary node present in slot
//                     E,N,W,S, 28-31 that boundary node carries a no-flux BC
//   tact[sty][stx]    1 if the 16x16 site tile was written by this step's spread
//                     (cleared by the interpolation after the lattice kernel)
// Mixed-precision contexts replace f/g with 16-bit deviations from the rest
// state (see lbib_lattice.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <deque>
#include <string>
#include <vector>
#include "../../include/lbib.h"

namespace lbib {

enum { DT = 0, DE = 1, DN = 2, DW = 3, DS = 4, DNE = 5, DNW = 6, DSW = 7, DSE = 8 };

constexpr uint32_t DOM_MASK = 0x00ffffffu;
constexpr int BN_SHIFT = 24;        // bit 24 + (slot-1), slot in E,N,W,S
constexpr int BC_SHIFT = 28;
constexpr int TILE_W = 256;         // lattice kernel block width = force tile width
constexpr int MAX_BC_SPECIES = LBIB_MAX_SPECIES;
constexpr int REINIT_LDS = 8192;          // level state of the changed sites kept in LDS
// Long connections of one remesh round ranked by a sort of their list; more
// of them fall back to a scan of the per-connection flags.
constexpr uint32_t RM_LIST_CAP = 4096;
constexpr int REINIT_CAP = 1 << 20;       // changed sites per step (global-memory state beyond REINIT_LDS)
constexpr int SEQ_PER_CONN = 256;   // candidates per connection (segment < 127 lu)
// Marker bins of the 16x16 spread tiles: BINCAP entries per tile, written in
// one pass by atomic slot counts; entries beyond a tile's capacity go to one
// shared overflow list of OVF_CAP entries that every reader also scans.
constexpr int BINCAP = 256;
constexpr uint32_t OVF_CAP = 1u << 16;
// Sharded counters: a statistic that thousands of workgroups add to is spread
// over NSHARD words 128 bytes apart (workgroup b adds to word b % NSHARD), so
// the atomics do not queue on one address; readers sum the shards.
constexpr int NSHARD = 64;
constexpr int SHARD_STRIDE = 32;    // uint32 words between shards (one 128-byte line)
enum { SH_TOUCHED = 0, SH_NBN = 1, SH_COUNT = 2 };

struct SpeciesDev {
    int kind;
    double tau, tauI;
    double production, decay, init_inside;
    uint32_t switchoff;
    int prod_domain;
    double source_inside;
    double krow[LBIB_MAX_SPECIES];
};

struct DevStatus {
    uint32_t bits;
    int32_t first_bad;
    uint32_t pool_used;     // unused
    uint32_t ncand;         // cut-link candidates of the current step
    uint32_t ncand_prev;    // candidates of the last completed cure (boundary-node reset list)
    uint32_t nchanged;      // sites whose local domain changed
    uint32_t nm, nc, maxid; // topology counts and the largest GeometryNode id
    uint32_t rm_L;          // long connections in the current remesh round
    uint32_t rm_total;      // connections split this step
    uint32_t co_any, co_removed;
    uint32_t co_nkeepc, co_nkeepm, co_nkeepr;
    uint32_t step_it;       // iteration of the last step begun with a clear status
    uint32_t rm_nc0, rm_nm0;   // counts before the current remesh round
    uint32_t reinit_levels;    // dependency levels of this step's CDE re-initialisation
    uint32_t rm_K;             // long connections listed by the current remesh round's flag pass
    uint32_t rm_sparse;        // the round's ranks come from the sorted list (else from the scan)
    uint32_t topo_changed;     // remesh or coarsening changed the topology in this step
    uint32_t novf;             // entries in the bin overflow list
    uint32_t cure_epoch;       // cures begun (tags the re-initialisation marks of the current one)
};

// Raised by remeshing for a new node within 2 sites of the edge; reported as
// LBIB_ERR_EDGE by lbib_read_status.  A bit of its own keeps it apart from the
// edge errors of the move and the cut links for the kernels below.
constexpr uint32_t ERR_EDGE_REMESH = 0x00000080u;
// A stopped status makes later kernels return without work.  Geometry and
// topology kernels stop on errors, halts and skipped steps; the lattice, IB,
// force and area-regulator kernels also while a remesh round is pending.
constexpr uint32_t STOP_BITS = LBIB_ERR_MASK | LBIB_HALT_MASK | LBIB_SKIPPED;
// The cut-link winners, the re-initialisation and the scanline run beside the
// remesh / ring / coarsening kernels of the same iteration, which the
// reference runs after them: they ignore what only those kernels raise.
constexpr uint32_t TOPO_ONLY_BITS = LBIB_HALT_PERTURB_REMESH | LBIB_HALT_GROW | LBIB_ERR_TOPOLOGY | ERR_EDGE_REMESH;
__device__ __forceinline__ bool halted_beside_topology(const DevStatus* st) {
    return (st->bits & (STOP_BITS & ~TOPO_ONLY_BITS)) != 0u;
}
constexpr uint32_t STOP_BITS_STEP = STOP_BITS | LBIB_FLAG_REMESH;
__device__ __forceinline__ bool step_stopped(const DevStatus* st) { return (st->bits & STOP_BITS_STEP) != 0u; }
// Opens an iteration: records it while the status is clear, otherwise marks
// the rest of the enqueued work as skipped (one thread).
// ITER_NEXT (a step recorded into a graph, replayed at many iterations):
// the iteration after the last one begun.
constexpr uint32_t ITER_NEXT = 0xffffffffu;
__device__ __forceinline__ void step_begin_dev(DevStatus* st, uint32_t it) {
    const uint32_t b = st->bits;
    if (b & STOP_BITS_STEP) st->bits = b | LBIB_SKIPPED;
    else st->step_it = (it == ITER_NEXT) ? st->step_it + 1u : it;
}
// the iteration a kernel argument names (ITER_NEXT: the one begun last)
__device__ __forceinline__ uint32_t iter_of(uint32_t it, const DevStatus* st) {
    return it == ITER_NEXT ? st->step_it : it;
}

// One periodic domain of a batched atlas (lbib_rect on the device).
struct RectDev {
    int32_t x0, y0, w, h;
    double tau, tauI;                        // fluid
    double tauIs[LBIB_MAX_SPECIES];          // species 1/tau
};

// The periodic frame of site (x, y): the whole lattice, or in a batched atlas
// the rectangle owning the site's 16x16 tile.  false: the site belongs to no
// domain and is never updated.
struct Frame {
    int x0, y0, w, h, r;
};
__device__ __forceinline__ bool site_frame(const int32_t* rtile, const RectDev* rects, int stx, int nx, int ny,
                                           int x, int y, Frame& fr) {
    if (rtile == nullptr) {
        fr.x0 = 0; fr.y0 = 0; fr.w = nx; fr.h = ny; fr.r = -1;
        return true;
    }
    const int r = rtile[(long)(y >> 4) * stx + (x >> 4)];
    if (r < 0) return false;
    const RectDev& R = rects[r];
    fr.x0 = R.x0; fr.y0 = R.y0; fr.w = R.w; fr.h = R.h; fr.r = r;
    return x - R.x0 < R.w && y - R.y0 < R.h;
}

// An entry of the bin overflow list: the tile it belongs to, the floor
// coordinates of the marker (ordered bins: floor(X) | floor(Y) << 16; push
// bins: those of the stencil corner), and either the marker index (ordered
// bins) or its offsets from the stencil corner and force in FP32 (push bins).
struct BinOvf {
    int32_t tile;
    uint32_t pos;
    int32_t k;
    float rx, ry, fx, fy;
};

// A cut-link candidate: its distance, its (site, slot) selection entry
// site * 4 + slot - 1, its global insertion number, the domain it carries and
// the connection's boundary descriptor (-1: none; the boundary-node velocity
// of a candidate with a descriptor is kept beside, in CandV).
struct Cand {
    double dist2;
    uint32_t ent;
    uint32_t seq;       // global insertion order: conn * SEQ_PER_CONN + k
    uint32_t domain;
    int32_t desc;
};
struct CandV {
    double vx, vy;
};
// Closest-wins selection entry of one (site, slot): the minimum distance (the
// bits of the double) and the minimum insertion number among candidates at
// that distance (kept only where two met at it); one 16-byte record, so a
// site's four slots share a 64-byte span.
struct SelEntry {
    unsigned long long dist;
    uint32_t seq;
    uint32_t pad;
};

}  // namespace lbib

struct lbib_ctx {
    lbib_config cfg;
    int nx = 0, ny = 0;
    long n = 0;
    int ns = 0;
    int prec = LBIB_FP64;
    int ntx = 0;                      // force tiles per row
    hipStream_t stream = nullptr;
    std::string err;
    size_t bytes = 0;
    // lattice
    void* f[2] = {nullptr, nullptr};
    void* g[2] = {nullptr, nullptr};
    int cur = 0;
    double *u = nullptr, *Fx = nullptr, *Fy = nullptr, *msite = nullptr;
    bool msite_live = false;
    uint32_t* tag = nullptr;
    lbib::SpeciesDev sp[LBIB_MAX_SPECIES];
    // batched atlas (lbib_set_batch_layout): rectangle per 16x16 tile
    int nrect = 0;
    int32_t* rtile = nullptr;
    lbib::RectDev* rects = nullptr;
    long live_sites = 0;               // sites inside some rectangle (all sites without a layout)
    bool rect_pairs = true;            // every rectangle width even: paired-lane kernel applies
    std::vector<int32_t> rect_host;    // x0, y0, w, h per rectangle
    void* vsegs = nullptr;             // vector lattice kernel: row segments of the atlas
    int vseg_len = 0, nvseg = 0;
    double* dconc = nullptr;           // per-domain concentration sums (lbib_domain_concentrations)
    int dconc_cap = 0;
    // lattice launch shape (lbib_autotune / lbib_set_lattice_shape)
    int lat_variant = -1, lat_wg = 256, lat_spt = 2;
    // spread over 16x16 tiles
    int stx = 0, sty = 0;
    // second stream: the topology stage (remesh, rings, coarsening) runs on it
    // beside the cut-link winners, re-initialisation and scanline, and the
    // next lattice stage's marker forces and tile binning follow it there.
    // topo_pending: the main stream has not yet waited for the topology stage
    // (ev_join); forces_beside: nothing enqueued since that stage changed what
    // the marker forces and binning read, so they may run on the second stream
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool topo_pending = false, forces_beside = false;
    // the marker bins are empty: the last kernel to touch them was the
    // interpolation, which clears them once the lattice kernel has read the
    // active-tile flags (allocation zero-fills them)
    bool bins_clear = true;
    uint32_t* tcnt = nullptr;          // [nt] entries binned into each tile (may exceed BINCAP)
    uint8_t* tact = nullptr;
    // ordered bins (FP64 spread, junction search): markers by the tile of
    // floor(X), floor(Y); BINCAP entries per tile
    int32_t* tlist = nullptr;          // [nt][BINCAP] marker index
    uint32_t* tpos = nullptr;          // [nt][BINCAP] floor(X) | floor(Y) << 16
    // push bins (mixed-precision spread): every tile a marker's stencil
    // reaches lists the stencil corner, the offsets from it and the force
    uint32_t* pb_pos = nullptr;        // [nt][BINCAP] x0 | y0 << 16
    float2 *pb_r = nullptr, *pb_f = nullptr;   // [nt][BINCAP] (X - x0, Y - y0), (Fx, Fy) in FP32
    lbib::BinOvf* ovf = nullptr;       // [OVF_CAP] entries past a tile's capacity (count: st->novf)
    uint32_t* shard = nullptr;         // [SH_COUNT][NSHARD][SHARD_STRIDE] sharded counters
    // markers
    int mcap = 0, nm = 0;
    double *mx = nullptr, *my = nullptr, *mvx = nullptr, *mvy = nullptr;
    double *mfx = nullptr, *mfy = nullptr, *mox = nullptr, *moy = nullptr;
    uint32_t* mid = nullptr;                   // GeometryNode id per dense index
    int32_t *min_c = nullptr, *mout_c = nullptr;   // incoming / outgoing connection
    double *sx_m = nullptr, *sy_m = nullptr, *svx_m = nullptr, *svy_m = nullptr, *sox_m = nullptr, *soy_m = nullptr;
    uint32_t* sid_m = nullptr;
    uint32_t *mdead = nullptr, *rec_alive = nullptr, *keep_m = nullptr, *pos_m = nullptr, *keep_r = nullptr,
             *pos_r = nullptr, *mdom = nullptr;
    int32_t *rec_a = nullptr, *rec_b = nullptr, *rec_bc = nullptr;
    // forces (CSR by marker)
    int fcap = 0;
    int32_t *fs_off = nullptr, *fs_type = nullptr, *fs_partner = nullptr;
    double* fs_prm = nullptr;
    // cell-junction rebuild (lbib_forces_junctions): per-marker closest
    // partner, incoming finders, new row counts/offsets, and a scratch CSR
    // of fcap structs the new table is assembled in
    bool junc_on = false;
    // device staging of the perturbation entry points (lbib_perturbation_*,
    // lbib_set_marker_positions), grown on demand
    char* pert_buf = nullptr;
    size_t pert_bytes = 0;
    // staging of lbib_add_mass_sites (site indices, then values)
    int64_t* mass_site = nullptr;
    double* mass_val = nullptr;
    long mass_cap = 0;
    int jcap = 0, jfcap = 0;
    int32_t *j_part = nullptr, *j_src = nullptr, *j_off = nullptr;
    uint32_t *j_in = nullptr, *j_cnt = nullptr, *j_ofs = nullptr, *j_total = nullptr;
    int32_t *j_type = nullptr, *j_partner = nullptr;
    double* j_prm = nullptr;
    // connections
    int ccap = 0, nc = 0;
    int32_t *ca = nullptr, *cb = nullptr, *cbc = nullptr;
    uint32_t* cdom = nullptr;
    int32_t *ca2 = nullptr, *cb2 = nullptr, *cbc2 = nullptr;   // compaction scratch
    uint32_t* cdom2 = nullptr;
    uint32_t *flag_c = nullptr, *rank_c = nullptr, *cdead = nullptr, *keep_c = nullptr, *pos_c = nullptr;
    uint32_t* ring_slot = nullptr;
    int32_t* ring_bucket = nullptr;
    int32_t* ring_marker = nullptr;
    uint32_t* scan_tmp = nullptr;
    uint32_t scan_cap = 0;
    uint32_t* rm_list = nullptr;       // a remesh round's long connections (RM_LIST_CAP entries)
    int bcd_cap = 0;
    int32_t *bcd_off = nullptr, *bcd_sp = nullptr;
    int nbcd = 0;
    bool any_bc = false;
    // cut links
    int candcap = 0;
    lbib::Cand* cand = nullptr;
    lbib::CandV* cand_v = nullptr;             // [candcap] (connections with a boundary condition)
    lbib::SelEntry* best = nullptr;            // [n][4] (site-major)
    uint32_t* bn_dom = nullptr;                // [4][n]
    double* bc_v = nullptr;                    // [4][n][2]
    int32_t* bc_desc = nullptr;                // [4][n]
    uint32_t* changed = nullptr;               // [REINIT_CAP][2]
    uint64_t* rmark = nullptr;                 // [n] (cure epoch << 32) | position in changed
    uint8_t* rstate = nullptr;                 // [REINIT_CAP] level state
    // rings / domains (indexed by domain id)
    int dcap = 0, ndom = 0;
    int32_t* ring_start = nullptr;
    uint32_t *ring_cnt = nullptr, *ring_off_u = nullptr;
    double* area = nullptr;
    uint32_t* celltype = nullptr;
    double* target = nullptr;
    double* mtab = nullptr;
    bool mtab_live = false;
    // measurement: event pairs around lattice-kernel launches
    int prof_every = 0;        // every prof_every-th lattice launch is timed (0 = none)
    uint64_t prof_seq = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    // run-ahead: host launch state when each iteration began
    struct Snapshot {
        uint32_t it;
        int cur;
        bool mtab_live, msite_live;
        size_t ev_used;
    };
    std::deque<Snapshot> snaps;
    // graph of a block of steps (lbib_capture_begin / _end, lbib_graph_launch)
    bool capturing = false;
    uint64_t alloc_gen = 0;               // bumped by every device allocation
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<Snapshot> gsnaps;         // launch state of each captured step (it relative)
    std::vector<int64_t> gkey;            // launch state the graph was recorded from
    // lbib_step_begin is carried out by the next kernel launched (the first
    // kernel of the lattice stage, or a kernel of its own)
    bool begin_pending = false;
    uint32_t begin_it = 0;
    // status
    lbib::DevStatus* st = nullptr;
    lbib::DevStatus* st_host = nullptr;   // pinned
    uint32_t* shard_host = nullptr;       // pinned copy of the sharded counters
};

namespace lbib {
// launchers implemented per translation unit
#define LBIB_HIP(call)                                                         \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            c->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return -1;                                                         \
        }                                                                      \
    } while (0)

int launch_init_f64(lbib_ctx* c, const double* u0_dev);
// the FP64 initial state into the given natural-order buffers (f[9][n],
// g[ns][5][n], velocity pairs u[n][2])
int launch_init_f64_into(lbib_ctx* c, const double* u0_dev, double* f, double* g, double* u);
int launch_stream_collide_f64(lbib_ctx* c, uint32_t it, bool use_ftile);
int launch_copy_probe(lbib_ctx* c, const void* src, void* dst, long nbytes);
int launch_init_mixed(lbib_ctx* c, const double* u0_dev);
int launch_stream_collide_mixed(lbib_ctx* c, uint32_t it, bool use_ftile);
// Lattice launch shape: variant 0 = wg-wide row tiles with spt sites per lane
// (FP64: spt 1), variant 1 = paired-lane 16-bit kernel, variant 2 = vector
// kernel with spt consecutive sites per lane (8- or 16-byte plane accesses).
struct LatticeShape {
    int variant, wg, spt;
};
// the paired-lane kernel needs 16-bit storage and even periodic widths
inline bool pairs_apply(const lbib_ctx* c) {
    return (c->prec == LBIB_FP16_DEV || c->prec == LBIB_BF16_DEV) && c->nx % 2 == 0 &&
           (c->nrect == 0 || c->rect_pairs);
}
// the vector kernel: mixed storage, v sites of 8 or 16 bytes per lane, rows
// made of whole lane vectors (rectangle origins are multiples of 16)
inline bool vec_apply(const lbib_ctx* c, int v) {
    if (c->prec == LBIB_FP64 || c->prec == LBIB_FP16_REM || c->nx % v) return false;
    if (c->prec == LBIB_FP32) return v == 2 || v == 4;
    return v == 4 || v == 8;
}
inline LatticeShape lattice_shape(const lbib_ctx* c) {
    if (c->lat_variant >= 0 && (c->lat_variant == 0 || (c->lat_variant == 1 && pairs_apply(c)) ||
                                (c->lat_variant == 2 && (vec_apply(c, c->lat_spt) || c->prec == LBIB_FP16_REM))))
        return {c->lat_variant, c->lat_wg, c->lat_spt};
    if (c->prec == LBIB_FP64 || c->prec == LBIB_FP16_REM) return {0, 256, 1};
    // one lattice: one site per lane in 256-wide groups.  With the
    // populations read non-temporally it beats the vector kernel (several
    // sites per lane, one wide load per plane) inside the coupled 2048^2
    // step in FP32 (91 against 97 us) and 16-bit storage (63 against 68 us);
    // a batched atlas keeps the paired or two-site shapes until the tuner
    // has timed them on its segments
    if (c->nrect == 0) return {0, 256, 1};
    return pairs_apply(c) ? LatticeShape{1, 256, 1} : LatticeShape{0, 256, 2};
}
struct TunedShape {
    LatticeShape shape;
    double ms;
};
// Times every applicable shape of the lattice kernel on the current state
// (reps launches each after one warm-up) into out.  The launches read the
// current buffers and write the next ones, a scratch velocity array and no
// mass or status, so the simulation is not advanced.
// Lattice-kernel tuner: one scratch velocity array and a clear status stand
// in for the live ones, everything else is the context's own state.
struct TuneScratch {
    void* u = nullptr;
    DevStatus* st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TuneScratch() {
        if (u) (void)hipFree(u);
        if (st) (void)hipFree(st);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    int init(lbib_ctx* c, size_t ubytes) {
        LBIB_HIP(hipMalloc(&u, ubytes));
        LBIB_HIP(hipMemcpyAsync(u, c->u, ubytes, hipMemcpyDeviceToDevice, c->stream));
        LBIB_HIP(hipMalloc((void**)&st, sizeof(DevStatus)));
        LBIB_HIP(hipMemsetAsync(st, 0, sizeof(DevStatus), c->stream));
        LBIB_HIP(hipEventCreate(&e0));
        LBIB_HIP(hipEventCreate(&e1));
        return 0;
    }
    template <typename Launch>
    int time(lbib_ctx* c, int reps, Launch launch, double& ms) {
        if (launch()) return -1;
        LBIB_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; r++)
            if (launch()) return -1;
        LBIB_HIP(hipEventRecord(e1, c->stream));
        LBIB_HIP(hipEventSynchronize(e1));
        float t = 0.0f;
        LBIB_HIP(hipEventElapsedTime(&t, e0, e1));
        ms = (double)t / reps;
        return 0;
    }
};
int autotune_f64(lbib_ctx* c, int reps, std::vector<TunedShape>& out);
int autotune_mixed(lbib_ctx* c, int reps, std::vector<TunedShape>& out);
int launch_force_and_spread(lbib_ctx* c);
int launch_clear_force(lbib_ctx* c);
int launch_interp_move(lbib_ctx* c);
int launch_geometry(lbib_ctx* c, uint32_t it, int mode);
int launch_area_regulator(lbib_ctx* c, double target, double kp, double maxs, double factor);
int scan_exclusive(lbib_ctx* c, const uint32_t* in, uint32_t* out, const uint32_t* nptr, uint32_t nfix, uint32_t cap,
                   uint32_t* total, const uint32_t* gate = nullptr);
int launch_inout(lbib_ctx* c);
// gated: skip (keeping the current ring order) unless st->topo_changed
int launch_rings(lbib_ctx* c, bool counters_cleared = false, bool gated = false);
// per-domain inputs and outputs of the area regulator's mass source
struct AreaMass {
    const uint32_t* celltype;
    double* target;
    double* mtab;
    double target_area, kp, maxs, factor;
};
// mass != nullptr: also the area regulator's mass table, skipped while the
// status is stopped (an iteration's BioSolver stage)
int launch_ring_areas(lbib_ctx* c, const AreaMass* mass);
// also clears the ring counters when the topology changed
int launch_remesh(lbib_ctx* c, int rounds);
// per-domain sum of one species' concentration into acc[ndom] (zeroed by the caller)
int launch_domain_concentrations(lbib_ctx* c, int species, int ndom, double* acc);
// a live context with any deferred step begin enqueued (-1: null context)
int ctx_ready(lbib_ctx* c);
int join_topology(lbib_ctx* c);
int launch_coarsen(lbib_ctx* c);
int launch_forces_clear(lbib_ctx* c);
// two structs of `type` per marker, one per adjacent connection (type 6:
// p0 = force; type 0: p0 = K, p1 = resting length)
int launch_forces_pairs(lbib_ctx* c, int type, double p0, double p1);
// per-site mass of tutorial_02_BioSolverGrowth from species sR and sL
int launch_mass_growth(lbib_ctx* c, int sR, int sL, double threshold, double maxsource);
int launch_mass_sites(lbib_ctx* c, long cnt, const int64_t* site, const double* val);
// the markers (of the first c->nm) whose position meets a perturbation
// condition of GeometryNode::perturbatePosition: *cnt (zeroed by the caller)
// counts them, the first cap get their dense index, id and position written
int launch_perturbation_candidates(lbib_ctx* c, int cap, uint32_t* cnt, int32_t* ix, uint32_t* id, double* xy);
// mx/my[ix[i]] = xy[2i], xy[2i+1] for i < n
int launch_set_positions(lbib_ctx* c, int n, const int32_t* ix, const double* xy);
// the domains (dflag[ndom] zeroed by the caller) holding a connection that
// meets the Connection::perturbConnection condition: *cnt counts them, the
// first cap are listed in dlist
int launch_conn_perturbation_domains(lbib_ctx* c, uint32_t* dflag, int cap, uint32_t* cnt, uint32_t* dlist);
// ring length and ring_marker offset of each of the nd listed domains
int launch_ring_info(lbib_ctx* c, int nd, const uint32_t* dlist, uint32_t* len, uint32_t* base);
// the markers of each listed ring in walk order from its first connection
// (node j is the first marker of ring connection j), at off[i]..off[i+1)
int launch_ring_gather(lbib_ctx* c, int nd, const uint32_t* base, const int32_t* off, int32_t* marker, double* xy);
// markers that may name one marker as their closest junction partner
constexpr int JUNCTION_MAXIN = 16;
// deleteForceType(7) and one type-7 pair per marker with a closest marker of
// another domain within radius (needs the junction scratch of ensure_junctions)
int launch_forces_junctions(lbib_ctx* c, double radius, double k, double l0);
// markers binned by the 16x16 tile of floor(X), floor(Y) (tcnt/toff/tlist)
int launch_tile_bins(lbib_ctx* c);
int launch_topology(lbib_ctx* c, bool coarsen);
// FP64 <-> storage conversion of the current buffers (mixed precision only)
int launch_pack(lbib_ctx* c, const double* src_dev, int field, int nq);
int launch_unpack(lbib_ctx* c, double* dst_dev, int field, int nq);
}  // namespace lbib

