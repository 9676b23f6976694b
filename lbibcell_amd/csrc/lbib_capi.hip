// lbib_capi.hip -- context management and the C ABI declared in include/lbib.h.
#include <algorithm>
#include <climits>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>
#include "lbib_internal.h"

using namespace lbib;

namespace lbib {
__global__ void k_step_begin(DevStatus* st, uint32_t it) { step_begin_dev(st, it); }

// the main stream waits for the topology stage still running on the second
// stream (its kernels raise stops the next iteration's begin must see)
// The join event is recorded on the second stream only when the main stream
// waits, after everything enqueued there so far (one marker packet fewer in
// the common case, where the forces and binning follow the topology stage).
int join_topology(lbib_ctx* c) {
    if (!c->topo_pending) return 0;
    c->topo_pending = false;
    LBIB_HIP(hipEventRecord(c->ev_join, c->stream2));
    LBIB_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return 0;
}
}  // namespace lbib

namespace {

template <typename T>
int dev_alloc(lbib_ctx* c, T** p, size_t count, int fill = 0) {
    *p = nullptr;
    if (c->capturing) {
        c->err = "device allocation while a step is being recorded into a graph";
        return -1;
    }
    c->alloc_gen++;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return -1;
    }
    c->bytes += count * sizeof(T);
    e = hipMemsetAsync(*p, fill, count * sizeof(T), c->stream);
    if (e != hipSuccess) {
        c->err = std::string("hipMemsetAsync: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}
template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

size_t storage_bytes(int prec) {
    switch (prec) {
        case LBIB_FP64: return 8;
        case LBIB_FP32: return 4;
        default: return 2;
    }
}
// bytes of one site's fluid (planes 9) or species (planes 5) populations
size_t field_bytes(int prec, int planes) {
    if (prec == LBIB_FP16_REM) return planes == 9 ? 3 * 4 + 9 * 2 : 4 + 5 * 2;
    return (size_t)planes * storage_bytes(prec);
}

int ensure_scan(lbib_ctx* c) {
    const uint32_t need =
        (uint32_t)((std::max({c->mcap, c->ccap, c->dcap, c->stx * c->sty, 1}) + 1023) / 1024 + 1);
    if (!c->rm_list && dev_alloc(c, &c->rm_list, RM_LIST_CAP)) return -1;
    if (need <= c->scan_cap) return 0;
    dev_free(c->scan_tmp);
    if (dev_alloc(c, &c->scan_tmp, need)) return -1;
    c->scan_cap = need;
    return 0;
}

int ensure_junctions(lbib_ctx* c);

// Marker-indexed arrays (positions, links, spread work lists, coarsening
// scratch, force CSR offsets).  Growth keeps every entry below the old
// capacity, so ids, links and the force offsets survive lbib_reserve and a
// set_markers call that crosses the capacity.
template <typename T>
static int regrow(lbib_ctx* c, T** p, size_t old_count, size_t count) {
    T* q = nullptr;
    if (dev_alloc(c, &q, count)) return -1;
    if (*p && old_count) {
        if (hipMemcpyAsync(q, *p, old_count * sizeof(T), hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            c->err = "copy of marker arrays failed";
            dev_free(q);
            return -1;
        }
    }
    dev_free(*p);
    *p = q;
    return 0;
}
int ensure_markers(lbib_ctx* c, int nm) {
    if (nm + 1 <= c->mcap) return 0;
    int cap = std::max(nm + 1, c->mcap * 2);
    cap = std::max(cap, 1024);
    const size_t old = (size_t)c->mcap;
    double** darr[14] = {&c->mx, &c->my, &c->mvx, &c->mvy, &c->mfx, &c->mfy, &c->mox, &c->moy,
                         &c->sx_m, &c->sy_m, &c->svx_m, &c->svy_m, &c->sox_m, &c->soy_m};
    uint32_t** uarr[9] = {&c->mid, &c->sid_m, &c->mdead, &c->rec_alive, &c->keep_m, &c->pos_m, &c->keep_r,
                          &c->pos_r, &c->mdom};
    int32_t** iarr[6] = {&c->min_c, &c->mout_c, &c->rec_a, &c->rec_b, &c->rec_bc, &c->fs_off};
    for (auto a : darr) if (regrow(c, a, old, cap)) return -1;
    for (auto a : uarr) if (regrow(c, a, old, cap)) return -1;
    for (auto a : iarr) if (regrow(c, a, old ? old + 1 : 0, cap + 1)) return -1;
    c->mcap = cap;
    if (ensure_scan(c)) return -1;
    return ensure_junctions(c);
}

// Connection-indexed arrays and the cut-link candidate list; growth keeps the
// entries below the old capacity (the connections and this step's candidates).
int ensure_connections(lbib_ctx* c, int nc) {
    if (nc <= c->ccap && c->cand) return 0;
    int cap = std::max(nc, c->ccap * 2);
    cap = std::max(cap, 1024);
    const size_t old = c->cand ? (size_t)c->ccap : 0;
    int32_t** iarr[8] = {&c->ca, &c->cb, &c->cbc, &c->ca2, &c->cb2, &c->cbc2, &c->ring_bucket, &c->ring_marker};
    uint32_t** uarr[8] = {&c->cdom, &c->cdom2, &c->flag_c, &c->rank_c, &c->cdead, &c->keep_c, &c->pos_c, &c->ring_slot};
    for (auto a : iarr) if (regrow(c, a, old, cap)) return -1;
    for (auto a : uarr) if (regrow(c, a, old, cap)) return -1;
    // a connection of length L crosses at most L+1 lines of each family; keep
    // room for 8 candidates per connection plus slack for long first segments.
    const int candcap = cap * 8 + 65536;
    if (regrow(c, &c->cand, c->cand ? (size_t)c->candcap : 0, candcap)) return -1;
    if (c->cand_v && regrow(c, &c->cand_v, (size_t)c->candcap, candcap)) return -1;
    c->candcap = candcap;
    c->ccap = cap;
    return ensure_scan(c);
}

int ensure_domains(lbib_ctx* c, int ndom) {
    if (ndom <= c->dcap) return 0;
    int cap = std::max(ndom, c->dcap * 2);
    cap = std::max(cap, 256);
    std::vector<double> tgt;
    if (c->target && c->dcap > 0) {
        tgt.resize(c->dcap);
        if (hipMemcpy(tgt.data(), c->target, c->dcap * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
            c->err = "copy of target areas failed";
            return -1;
        }
    }
    dev_free(c->celltype); dev_free(c->target); dev_free(c->mtab); dev_free(c->ring_start); dev_free(c->ring_cnt);
    dev_free(c->ring_off_u); dev_free(c->area);
    if (dev_alloc(c, &c->celltype, cap) || dev_alloc(c, &c->target, cap, 0xff) || dev_alloc(c, &c->mtab, cap) ||
        dev_alloc(c, &c->ring_start, cap) || dev_alloc(c, &c->ring_cnt, cap) || dev_alloc(c, &c->ring_off_u, cap) ||
        dev_alloc(c, &c->area, cap))
        return -1;
    if (!tgt.empty() && hipMemcpy(c->target, tgt.data(), tgt.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        c->err = "copy of target areas failed";
        return -1;
    }
    c->dcap = cap;
    return ensure_scan(c);
}

// a pending lbib_step_begin that no lattice stage consumed gets a kernel of
// its own before anything else is enqueued
int flush_step_begin(lbib_ctx* c) {
    if (!c->begin_pending) return 0;
    c->begin_pending = false;
    k_step_begin<<<1, 1, 0, c->stream>>>(c->st, c->capturing ? ITER_NEXT : c->begin_it);
    LBIB_HIP(hipGetLastError());
    return 0;
}
// Every entry point: later work sees the topology stage and the pending step
// begin; the marker forces of the next lattice stage then run on the main
// stream (the call may change what they read).
int check_ctx(lbib_ctx* c) {
    if (!c) return -1;
    c->forces_beside = false;
    if (join_topology(c)) return -1;
    return flush_step_begin(c);
}
// Entry points that change nothing the marker forces and binning read (the
// area regulator, the step begin) keep them on the second stream.
int check_ctx_keep_forces(lbib_ctx* c) {
    if (!c) return -1;
    if (join_topology(c)) return -1;
    return flush_step_begin(c);
}
int check_ctx_keep_begin(lbib_ctx* c) { return c ? 0 : -1; }
int upload(lbib_ctx* c, void* dst, const void* src, size_t bytes);
int download(lbib_ctx* c, void* dst, const void* src, size_t bytes);

// ensure_connections that keeps the previous cure's candidate list (the next
// stage B clears those boundary-node bits) valid across the reallocation
int grow_connections(lbib_ctx* c, int nc) { return ensure_connections(c, nc); }

// Force structs: growth keeps the table's contents (a BioSolver may add to
// structs written earlier in the same iteration).
int ensure_force_cap(lbib_ctx* c, int nf) {
    if (nf <= c->fcap) return 0;
    const int old = c->fcap;
    const int cap = std::max(nf, 2 * c->fcap);
    int32_t *t = nullptr, *p = nullptr;
    double* q = nullptr;
    if (dev_alloc(c, &t, cap) || dev_alloc(c, &p, cap) || dev_alloc(c, &q, 4 * (size_t)cap)) {
        dev_free(t); dev_free(p); dev_free(q);
        return -1;
    }
    if (old > 0) {
        LBIB_HIP(hipMemcpyAsync(t, c->fs_type, old * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
        LBIB_HIP(hipMemcpyAsync(p, c->fs_partner, old * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
        LBIB_HIP(hipMemcpyAsync(q, c->fs_prm, 4 * (size_t)old * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        LBIB_HIP(hipStreamSynchronize(c->stream));
    }
    dev_free(c->fs_type); dev_free(c->fs_partner); dev_free(c->fs_prm);
    c->fs_type = t; c->fs_partner = p; c->fs_prm = q;
    c->fcap = cap;
    if (c->junc_on && c->jfcap < c->fcap) {
        dev_free(c->j_type); dev_free(c->j_partner); dev_free(c->j_prm);
        if (dev_alloc(c, &c->j_type, c->fcap) || dev_alloc(c, &c->j_partner, c->fcap) ||
            dev_alloc(c, &c->j_prm, 4 * (size_t)c->fcap))
            return -1;
        c->jfcap = c->fcap;
    }
    return 0;
}

// Scratch of the cell-junction rebuild, sized with the markers once the
// junction BioSolver has run: later growth happens with the marker arrays,
// never inside a step recorded into a graph.  The force table keeps room for
// two tension structs, a junction and an average of one incoming junction per
// marker; a fuller table is an LBIB_ERR_CAPACITY at run time.
int ensure_junctions(lbib_ctx* c) {
    if (!c->junc_on) return 0;
    if (c->jcap < c->mcap) {
        int32_t** iarr[3] = {&c->j_part, &c->j_src, &c->j_off};
        uint32_t** uarr[3] = {&c->j_in, &c->j_cnt, &c->j_ofs};
        for (auto a : iarr) dev_free(*a);
        for (auto a : uarr) dev_free(*a);
        if (dev_alloc(c, &c->j_part, c->mcap) || dev_alloc(c, &c->j_src, (size_t)c->mcap * JUNCTION_MAXIN) ||
            dev_alloc(c, &c->j_off, c->mcap + 1) || dev_alloc(c, &c->j_in, c->mcap) ||
            dev_alloc(c, &c->j_cnt, c->mcap) || dev_alloc(c, &c->j_ofs, c->mcap))
            return -1;
        c->jcap = c->mcap;
    }
    if (!c->j_total && dev_alloc(c, &c->j_total, 1)) return -1;
    if (ensure_force_cap(c, 4 * c->mcap)) return -1;
    if (c->jfcap < c->fcap) {
        dev_free(c->j_type); dev_free(c->j_partner); dev_free(c->j_prm);
        if (dev_alloc(c, &c->j_type, c->fcap) || dev_alloc(c, &c->j_partner, c->fcap) ||
            dev_alloc(c, &c->j_prm, 4 * (size_t)c->fcap))
            return -1;
        c->jfcap = c->fcap;
    }
    return 0;
}

// write the host's marker/connection counts into the device status
int put_counts(lbib_ctx* c) {
    uint32_t v[2] = {(uint32_t)c->nm, (uint32_t)c->nc};
    static_assert(offsetof(DevStatus, nc) == offsetof(DevStatus, nm) + sizeof(uint32_t), "layout");
    LBIB_HIP(hipMemcpyAsync((char*)c->st + offsetof(DevStatus, nm), v, sizeof(v), hipMemcpyHostToDevice, c->stream));
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int upload(lbib_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return 0;
    LBIB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return 0;
}
int download(lbib_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return 0;
    LBIB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// Stage a host FP64 field into device scratch (the force arrays are reused as
// staging only when no step is in flight; a dedicated buffer keeps it simple).
struct Staging {
    double* p = nullptr;
    ~Staging() { if (p) (void)hipFree(p); }
};

// host launch state that the recorded kernel arguments and grids depend on
std::vector<int64_t> graph_key(const lbib_ctx* c) {
    const LatticeShape sh = lattice_shape(c);
    return {(int64_t)c->alloc_gen, c->cur, c->mtab_live, c->msite_live, c->nm == 0, c->nc == 0, c->ndom,
            sh.variant, sh.wg, sh.spt, c->prof_every, c->nrect, c->bins_clear};
}
void graph_drop(lbib_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->gsnaps.clear();
    c->gkey.clear();
}
}  // namespace

namespace lbib {
int ctx_ready(lbib_ctx* c) { return check_ctx(c); }
}  // namespace lbib

extern "C" {

int lbib_create(const lbib_config* cfg, lbib_ctx** out) {
    if (!cfg || !out) return -1;
    *out = nullptr;
    lbib_ctx* c = new (std::nothrow) lbib_ctx();
    if (!c) return -1;
    c->cfg = *cfg;
    // the cut-link selection entries index (site, slot) pairs in 32 bits
    if (cfg->nx < 4 || cfg->ny < 4 || cfg->nx > 65535 || cfg->ny > 65535 ||
        (long)cfg->nx * (long)cfg->ny > (1L << 30) || cfg->nspecies < 0 || cfg->nspecies > LBIB_MAX_SPECIES ||
        !(cfg->tau_fluid > 0.5)) {
        c->err = "invalid configuration (4 <= nx, ny <= 65535, nx * ny <= 2^30, 0 <= nspecies <= 4, "
                 "tau_fluid > 0.5)";
        *out = c;
        return -1;
    }
    if (cfg->precision != LBIB_FP64 && cfg->precision != LBIB_FP32 && cfg->precision != LBIB_FP16_DEV &&
        cfg->precision != LBIB_BF16_DEV && cfg->precision != LBIB_FP16_REM) {
        c->err = "unknown precision";
        *out = c;
        return -1;
    }
    for (int s = 0; s < cfg->nspecies; s++) {
        const lbib_species_spec& p = cfg->species[s];
        if (p.kind < LBIB_SPECIES_PLAIN || p.kind > LBIB_SPECIES_SCHNAKENBERG_L || !(p.tau > 0.0)) {
            c->err = "species " + std::to_string(s) + ": unknown kind or tau <= 0";
            *out = c;
            return -1;
        }
        const bool sch = p.kind == LBIB_SPECIES_SCHNAKENBERG_R || p.kind == LBIB_SPECIES_SCHNAKENBERG_L;
        if (sch && (p.prod_domain < 0 || p.prod_domain >= cfg->nspecies || p.prod_domain == s)) {
            c->err = "species " + std::to_string(s) + ": the Schnakenberg partner index must name another species";
            *out = c;
            return -1;
        }
    }
    *out = c;
    c->nx = cfg->nx;
    c->ny = cfg->ny;
    c->n = (long)cfg->nx * cfg->ny;
    c->ns = cfg->nspecies;
    c->prec = cfg->precision;
    c->ntx = (c->nx + TILE_W - 1) / TILE_W;
    for (int s = 0; s < LBIB_MAX_SPECIES; s++) {
        SpeciesDev& d = c->sp[s];
        memset(&d, 0, sizeof(d));
        if (s < c->ns) {
            const lbib_species_spec& p = cfg->species[s];
            d.kind = p.kind;
            d.tau = p.tau;
            d.tauI = 1.0 / p.tau;
            d.production = p.production;
            d.decay = p.decay;
            d.init_inside = p.init_inside;
            d.switchoff = p.switchoff;
            d.prod_domain = p.prod_domain;
            d.source_inside = p.source_inside;
            for (int j = 0; j < LBIB_MAX_SPECIES; j++) d.krow[j] = p.krow[j];
        }
    }
    LBIB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    LBIB_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    // the fork and join only order work on this device: a device-scope
    // release is enough (the default system-scope fence writes back caches)
    LBIB_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming | hipEventReleaseToDevice));
    LBIB_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming | hipEventReleaseToDevice));
    const long n = c->n;
    const long npad = (n + 63) / 64 * 64;   // whole 64-site blocks for 32- and 16-bit storage
    for (int b = 0; b < 2; b++) {
        if (dev_alloc(c, (uint8_t**)&c->f[b], npad * field_bytes(c->prec, 9))) return -1;
        if (dev_alloc(c, (uint8_t**)&c->g[b], (size_t)std::max(c->ns, 1) * npad * field_bytes(c->prec, 5))) return -1;
    }
    if (dev_alloc(c, &c->u, 2 * n) || dev_alloc(c, &c->Fx, n) || dev_alloc(c, &c->Fy, n) ||
        dev_alloc(c, &c->msite, n))
        return -1;
    c->stx = (c->nx + 15) / 16;
    c->sty = (c->ny + 15) / 16;
    c->live_sites = c->n;
    {
        const size_t nt = (size_t)c->stx * c->sty;
        if (dev_alloc(c, &c->tag, n) || dev_alloc(c, &c->tact, nt) || dev_alloc(c, &c->tcnt, nt))
            return -1;
        // marker bins: BINCAP entries per tile (ordered bins always, push
        // bins for the order-free spread of mixed-precision lattices)
        const size_t nb = nt * (size_t)BINCAP;
        if (dev_alloc(c, &c->tlist, nb) || dev_alloc(c, &c->tpos, nb) || dev_alloc(c, &c->ovf, (size_t)OVF_CAP) ||
            dev_alloc(c, &c->shard, (size_t)SH_COUNT * NSHARD * SHARD_STRIDE))
            return -1;
        if (c->prec != LBIB_FP64 &&
            (dev_alloc(c, &c->pb_pos, nb) || dev_alloc(c, &c->pb_r, nb) || dev_alloc(c, &c->pb_f, nb)))
            return -1;
    }
    if (dev_alloc(c, &c->best, 4 * n, 0xff) || dev_alloc(c, &c->bn_dom, 4 * n)) return -1;
    if (dev_alloc(c, &c->changed, 2 * (size_t)REINIT_CAP) || dev_alloc(c, &c->rmark, (size_t)c->n) ||
        dev_alloc(c, &c->rstate, (size_t)REINIT_CAP))
        return -1;
    if (dev_alloc(c, &c->st, 1)) return -1;
    LBIB_HIP(hipHostMalloc((void**)&c->st_host, sizeof(DevStatus), hipHostMallocDefault));
    LBIB_HIP(hipHostMalloc((void**)&c->shard_host, (size_t)SH_COUNT * NSHARD * SHARD_STRIDE * sizeof(uint32_t),
                           hipHostMallocDefault));
    if (ensure_markers(c, cfg->max_markers > 0 ? cfg->max_markers : 1024)) return -1;
    if (ensure_connections(c, cfg->max_connections > 0 ? cfg->max_connections : 1024)) return -1;
    if (ensure_domains(c, 256)) return -1;
    // empty force CSR
    if (dev_alloc(c, &c->fs_type, 1) || dev_alloc(c, &c->fs_partner, 1) || dev_alloc(c, &c->fs_prm, 4)) return -1;
    c->fcap = 1;
    if (put_counts(c)) return -1;
    {
        int32_t one[2] = {0, 0};
        if (dev_alloc(c, &c->bcd_off, 2) || dev_alloc(c, &c->bcd_sp, 1)) return -1;
        if (upload(c, c->bcd_off, one, sizeof(one))) return -1;
    }
    {
        DevStatus s0;
        memset(&s0, 0, sizeof(s0));
        s0.first_bad = INT_MAX;
        if (upload(c, c->st, &s0, sizeof(s0))) return -1;
    }
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

void lbib_destroy(lbib_ctx* c) {
    if (!c) return;
    if (c->capturing) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c->capturing = false;
    }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int b = 0; b < 2; b++) { dev_free(c->f[b]); dev_free(c->g[b]); }
    dev_free(c->u); dev_free(c->Fx); dev_free(c->Fy); dev_free(c->msite);
    dev_free(c->mass_site); dev_free(c->mass_val); dev_free(c->pert_buf);
    dev_free(c->rtile); dev_free(c->rects); dev_free(c->dconc); dev_free(c->vsegs);
    dev_free(c->tag); dev_free(c->tact); dev_free(c->tcnt);
    dev_free(c->tlist); dev_free(c->tpos); dev_free(c->ovf); dev_free(c->shard);
    dev_free(c->pb_pos); dev_free(c->pb_r); dev_free(c->pb_f);
    double** darr[14] = {&c->mx, &c->my, &c->mvx, &c->mvy, &c->mfx, &c->mfy, &c->mox, &c->moy,
                         &c->sx_m, &c->sy_m, &c->svx_m, &c->svy_m, &c->sox_m, &c->soy_m};
    uint32_t** uarr[19] = {&c->mid, &c->sid_m, &c->mdead, &c->rec_alive, &c->keep_m, &c->pos_m, &c->keep_r,
                           &c->pos_r, &c->mdom, &c->cdom, &c->cdom2, &c->flag_c, &c->rank_c, &c->cdead, &c->keep_c,
                           &c->pos_c, &c->ring_slot, &c->scan_tmp, &c->rm_list};
    int32_t** iarr[16] = {&c->min_c, &c->mout_c, &c->rec_a, &c->rec_b, &c->rec_bc, &c->fs_off, &c->ca, &c->cb,
                          &c->cbc, &c->ca2, &c->cb2, &c->cbc2, &c->ring_bucket, &c->ring_marker, &c->fs_type,
                          &c->fs_partner};
    for (auto a : darr) dev_free(*a);
    for (auto a : uarr) dev_free(*a);
    for (auto a : iarr) dev_free(*a);
    dev_free(c->fs_prm); dev_free(c->bcd_off); dev_free(c->bcd_sp);
    dev_free(c->j_part); dev_free(c->j_src); dev_free(c->j_off); dev_free(c->j_in); dev_free(c->j_cnt);
    dev_free(c->j_ofs); dev_free(c->j_total); dev_free(c->j_type); dev_free(c->j_partner); dev_free(c->j_prm);
    dev_free(c->cand); dev_free(c->cand_v); dev_free(c->best); dev_free(c->bn_dom); dev_free(c->bc_v);
    dev_free(c->bc_desc); dev_free(c->changed); dev_free(c->rmark); dev_free(c->rstate);
    dev_free(c->ring_start); dev_free(c->ring_cnt); dev_free(c->ring_off_u); dev_free(c->area);
    dev_free(c->celltype); dev_free(c->target); dev_free(c->mtab); dev_free(c->st);
    if (c->st_host) (void)hipHostFree(c->st_host);
    if (c->shard_host) (void)hipHostFree(c->shard_host);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* lbib_last_error(const lbib_ctx* c) { return c ? c->err.c_str() : "null context"; }

int lbib_synchronize(lbib_ctx* c) {
    if (check_ctx(c)) return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

size_t lbib_device_bytes(const lbib_ctx* c) { return c ? c->bytes : 0; }

// ---------------------------------------------------------------- lattice --
int lbib_init_solvers(lbib_ctx* c, const double* u0) {
    if (check_ctx(c)) return -1;
    Staging st;
    if (u0) {
        LBIB_HIP(hipMalloc((void**)&st.p, 2 * c->n * sizeof(double)));
        if (upload(c, st.p, u0, 2 * c->n * sizeof(double))) return -1;
    }
    int r = (c->prec == LBIB_FP64) ? launch_init_f64(c, st.p) : launch_init_mixed(c, st.p);
    if (r) return r;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

static int field_io(lbib_ctx* c, double* host, int field, int nq, bool to_device) {
    // field: 0 = fluid, 1 + s = species s
    void* base = field == 0 ? c->f[c->cur] : c->g[c->cur];
    const size_t off = field == 0 ? 0 : (size_t)(field - 1) * 5 * c->n;
    const size_t count = (size_t)nq * c->n;
    if (c->prec == LBIB_FP64) {
        double* d = (double*)base + off;
        if (to_device) return upload(c, d, host, count * sizeof(double)) ? -1 : (hipStreamSynchronize(c->stream) == hipSuccess ? 0 : -1);
        return download(c, host, d, count * sizeof(double));
    }
    Staging st;
    LBIB_HIP(hipMalloc((void**)&st.p, count * sizeof(double)));
    if (to_device) {
        if (upload(c, st.p, host, count * sizeof(double))) return -1;
        if (launch_pack(c, st.p, field, nq)) return -1;
        LBIB_HIP(hipStreamSynchronize(c->stream));
        return 0;
    }
    if (launch_unpack(c, st.p, field, nq)) return -1;
    return download(c, host, st.p, count * sizeof(double));
}

int lbib_set_populations(lbib_ctx* c, const double* f) {
    if (check_ctx(c) || !f) return -1;
    return field_io(c, const_cast<double*>(f), 0, 9, true);
}
int lbib_get_populations(lbib_ctx* c, double* f) {
    if (check_ctx(c) || !f) return -1;
    return field_io(c, f, 0, 9, false);
}
int lbib_set_species(lbib_ctx* c, int s, const double* g) {
    if (check_ctx(c) || !g) return -1;
    if (s < 0 || s >= c->ns) { c->err = "species index out of range"; return -1; }
    return field_io(c, const_cast<double*>(g), 1 + s, 5, true);
}
int lbib_get_species(lbib_ctx* c, int s, double* g) {
    if (check_ctx(c) || !g) return -1;
    if (s < 0 || s >= c->ns) { c->err = "species index out of range"; return -1; }
    return field_io(c, g, 1 + s, 5, false);
}
int lbib_get_velocity(lbib_ctx* c, double* ux, double* uy) {
    if (check_ctx(c) || !ux || !uy) return -1;
    if (c->prec == LBIB_FP64) {
        std::vector<double> u(2 * (size_t)c->n);
        if (download(c, u.data(), c->u, u.size() * sizeof(double))) return -1;
        for (long i = 0; i < c->n; i++) {
            ux[i] = u[2 * i];
            uy[i] = u[2 * i + 1];
        }
        return 0;
    }
    // mixed precision keeps FP32 pairs in the same allocation
    std::vector<float> u(2 * (size_t)c->n);
    if (download(c, u.data(), c->u, u.size() * sizeof(float))) return -1;
    for (long i = 0; i < c->n; i++) {
        ux[i] = (double)u[2 * i];
        uy[i] = (double)u[2 * i + 1];
    }
    return 0;
}
int lbib_get_domains(lbib_ctx* c, uint32_t* dom) {
    if (check_ctx(c) || !dom) return -1;
    if (download(c, dom, c->tag, c->n * sizeof(uint32_t))) return -1;
    for (long i = 0; i < c->n; i++) dom[i] &= DOM_MASK;
    return 0;
}
int lbib_set_domains(lbib_ctx* c, const uint32_t* dom) {
    if (check_ctx(c) || !dom) return -1;
    std::vector<uint32_t> t(c->n);
    if (download(c, t.data(), c->tag, c->n * sizeof(uint32_t))) return -1;
    for (long i = 0; i < c->n; i++) {
        if (dom[i] > DOM_MASK) { c->err = "domain id exceeds 24 bits"; return -1; }
        t[i] = (t[i] & ~DOM_MASK) | dom[i];
    }
    if (upload(c, c->tag, t.data(), c->n * sizeof(uint32_t))) return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
// Mixed-precision contexts keep the force as FP32 (Fx, Fy) pairs in Fx's
// allocation.
int lbib_set_force_field(lbib_ctx* c, const double* fx, const double* fy) {
    if (check_ctx(c) || !fx || !fy) return -1;
    if (c->prec == LBIB_FP64) {
        if (upload(c, c->Fx, fx, c->n * sizeof(double)) || upload(c, c->Fy, fy, c->n * sizeof(double))) return -1;
    } else {
        std::vector<float> p(2 * (size_t)c->n);
        for (long i = 0; i < c->n; i++) {
            p[2 * i] = (float)fx[i];
            p[2 * i + 1] = (float)fy[i];
        }
        if (upload(c, c->Fx, p.data(), p.size() * sizeof(float))) return -1;
    }
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
int lbib_get_force_field(lbib_ctx* c, double* fx, double* fy) {
    if (check_ctx(c) || !fx || !fy) return -1;
    if (c->prec == LBIB_FP64) {
        if (download(c, fx, c->Fx, c->n * sizeof(double))) return -1;
        return download(c, fy, c->Fy, c->n * sizeof(double));
    }
    std::vector<float> p(2 * (size_t)c->n);
    if (download(c, p.data(), c->Fx, p.size() * sizeof(float))) return -1;
    for (long i = 0; i < c->n; i++) {
        fx[i] = (double)p[2 * i];
        fy[i] = (double)p[2 * i + 1];
    }
    return 0;
}
int lbib_set_mass_field(lbib_ctx* c, const double* m) {
    if (check_ctx(c) || !m) return -1;
    if (upload(c, c->msite, m, c->n * sizeof(double))) return -1;
    c->msite_live = true;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------- markers --
// GeometryHandler::connectGeometryNodesToPhysicalNodes (GeometryHandler.cpp:
// 613-636) fails for a node whose 4x4 stencil leaves the lattice, and every
// node is connected when it is created.  The same bound, 2 <= X <= nx - 2 and
// 2 <= Y <= ny - 2 (inside the owning rectangle in a batched atlas), is
// required of every uploaded position before anything is launched: the spread
// tiles and the interpolation index the lattice without wrapping.
static int check_marker_edges(lbib_ctx* c, int nm, const double* xy, const int32_t* index = nullptr) {
    std::vector<int32_t> tmap;
    const int stx = (c->nx + 15) / 16, sty = (c->ny + 15) / 16;
    if (c->nrect > 0) {
        tmap.assign((size_t)stx * sty, -1);
        for (int r = 0; r < c->nrect; r++) {
            const int32_t* q = &c->rect_host[4 * r];
            for (int ty = q[1] / 16; ty < (q[1] + q[3] + 15) / 16 && ty < sty; ty++)
                for (int tx = q[0] / 16; tx < (q[0] + q[2] + 15) / 16 && tx < stx; tx++) tmap[(size_t)ty * stx + tx] = r;
        }
    }
    for (int k = 0; k < nm; k++) {
        const double X = xy[2 * k], Y = xy[2 * k + 1];
        double x0 = 0, y0 = 0, w = c->nx, h = c->ny;
        bool ok = std::isfinite(X) && std::isfinite(Y) && X >= 0 && Y >= 0 && X < c->nx && Y < c->ny;
        if (ok && c->nrect > 0) {
            const int r = tmap[(size_t)((int)Y / 16) * stx + (int)X / 16];
            if (r < 0) {
                ok = false;
            } else {
                const int32_t* q = &c->rect_host[4 * r];
                x0 = q[0]; y0 = q[1]; w = q[2]; h = q[3];
            }
        }
        if (ok) ok = X - 2.0 >= x0 && Y - 2.0 >= y0 && X + 2.0 <= x0 + w && Y + 2.0 <= y0 + h;
        if (!ok) {
            const int32_t at = index ? index[k] : k;
            struct { uint32_t bits; int32_t first; } z = {LBIB_ERR_EDGE, at};
            LBIB_HIP(hipMemcpyAsync(c->st, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
            LBIB_HIP(hipStreamSynchronize(c->stream));
            char msg[192];
            snprintf(msg, sizeof msg, "geometry point %d at (%.17g, %.17g) lies within 2 sites of the edge of its domain",
                     at, X, Y);
            c->err = msg;
            return -1;
        }
    }
    return 0;
}

int lbib_set_markers(lbib_ctx* c, int nm, const double* xy, const double* vxy, const double* oxy) {
    if (check_ctx(c) || nm < 0 || (nm > 0 && !xy)) return -1;
    if (check_marker_edges(c, nm, xy)) return -1;
    if (ensure_markers(c, nm)) return -1;
    std::vector<double> a(nm), b(nm);
    auto split = [&](const double* src, double* dx, double* dy) -> int {
        for (int k = 0; k < nm; k++) { a[k] = src[2 * k]; b[k] = src[2 * k + 1]; }
        if (upload(c, dx, a.data(), nm * sizeof(double)) || upload(c, dy, b.data(), nm * sizeof(double))) return -1;
        LBIB_HIP(hipStreamSynchronize(c->stream));
        return 0;
    };
    if (nm > 0) {
        if (split(xy, c->mx, c->my)) return -1;
        if (vxy && split(vxy, c->mvx, c->mvy)) return -1;
        if (oxy && split(oxy, c->mox, c->moy)) return -1;
    }
    c->nm = nm;
    return put_counts(c);
}
// the device marker and connection counts, read after everything enqueued
static int sync_counts(lbib_ctx* c) {
    LBIB_HIP(hipMemcpyAsync(c->st_host, c->st, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream));
    LBIB_HIP(hipStreamSynchronize(c->stream));
    c->nm = (int)c->st_host->nm;
    c->nc = (int)c->st_host->nc;
    return 0;
}
// Carves typed, 256-byte aligned arrays out of the perturbation staging: a
// first pass with base == nullptr sizes it.
extern "C++" {
struct Carve {
    char* base;
    size_t at = 0;
    template <typename T>
    T* take(size_t n) {
        at = (at + 255) & ~size_t(255);
        T* r = reinterpret_cast<T*>(base + at);
        at += n * sizeof(T);
        return r;
    }
};
}
static char* pert_staging(lbib_ctx* c, size_t bytes) {
    if (bytes > c->pert_bytes) {
        dev_free(c->pert_buf);
        c->pert_bytes = 0;
        if (dev_alloc(c, &c->pert_buf, bytes)) return nullptr;
        c->pert_bytes = bytes;
    }
    return c->pert_buf;
}
int lbib_perturbation_candidates(lbib_ctx* c, int cap, int* count, int32_t* index, uint32_t* id, double* xy) {
    if (check_ctx(c) || cap < 0 || !count || (cap > 0 && (!index || !id || !xy))) return -1;
    if (sync_counts(c)) return -1;
    const size_t w = (size_t)std::max(cap, 1);
    Carve sz{nullptr};
    sz.take<uint32_t>(1); sz.take<double>(2 * w); sz.take<int32_t>(w); sz.take<uint32_t>(w);
    Carve cv{pert_staging(c, sz.at)};
    if (!cv.base) return -1;
    uint32_t* cnt = cv.take<uint32_t>(1);
    double* dxy = cv.take<double>(2 * w);
    int32_t* dix = cv.take<int32_t>(w);
    uint32_t* did = cv.take<uint32_t>(w);
    LBIB_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), c->stream));
    if (launch_perturbation_candidates(c, cap, cnt, dix, did, dxy)) return -1;
    uint32_t n = 0;
    if (download(c, &n, cnt, sizeof n)) return -1;
    *count = (int)n;
    const size_t got = std::min<size_t>(n, (size_t)cap);
    if (got > 0 && (download(c, index, dix, got * sizeof(int32_t)) || download(c, id, did, got * sizeof(uint32_t)) ||
                    download(c, xy, dxy, 2 * got * sizeof(double))))
        return -1;
    return 0;
}
int lbib_perturbation_rings(lbib_ctx* c, int cap_rings, int cap_nodes, int* nrings, int* nnodes, uint32_t* domain,
                            int32_t* off, int32_t* marker, double* xy) {
    if (check_ctx(c) || cap_rings < 0 || cap_nodes < 0 || !nrings || !nnodes ||
        (cap_rings > 0 && (!domain || !off)) || (cap_nodes > 0 && (!marker || !xy)))
        return -1;
    *nrings = 0;
    *nnodes = 0;
    if (sync_counts(c)) return -1;
    if (c->nc == 0 || c->ndom == 0) return 0;
    const size_t wr = (size_t)std::max(cap_rings, 1), wn = (size_t)std::max(cap_nodes, 1);
    Carve sz{nullptr};
    sz.take<uint32_t>(c->ndom); sz.take<uint32_t>(1); sz.take<uint32_t>(wr); sz.take<uint32_t>(wr);
    sz.take<uint32_t>(wr); sz.take<int32_t>(wr + 1); sz.take<int32_t>(wn); sz.take<double>(2 * wn);
    Carve cv{pert_staging(c, sz.at)};
    if (!cv.base) return -1;
    uint32_t* dflag = cv.take<uint32_t>(c->ndom);
    uint32_t* cnt = cv.take<uint32_t>(1);
    uint32_t* dlist = cv.take<uint32_t>(wr);
    uint32_t* dlen = cv.take<uint32_t>(wr);
    uint32_t* dbase = cv.take<uint32_t>(wr);
    int32_t* doff = cv.take<int32_t>(wr + 1);
    int32_t* dmarker = cv.take<int32_t>(wn);
    double* dxy = cv.take<double>(2 * wn);
    // ring order of the current topology (the in/out links first)
    if (launch_inout(c) || launch_rings(c)) return -1;
    LBIB_HIP(hipMemsetAsync(dflag, 0, (size_t)c->ndom * sizeof(uint32_t), c->stream));
    LBIB_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), c->stream));
    if (launch_conn_perturbation_domains(c, dflag, cap_rings, cnt, dlist)) return -1;
    uint32_t nd = 0;
    if (download(c, &nd, cnt, sizeof nd)) return -1;
    *nrings = (int)nd;
    if (nd == 0 || nd > (uint32_t)cap_rings) return 0;
    std::vector<uint32_t> doms(nd), len(nd), base(nd);
    if (download(c, doms.data(), dlist, nd * sizeof(uint32_t))) return -1;
    std::sort(doms.begin(), doms.end());
    if (upload(c, dlist, doms.data(), nd * sizeof(uint32_t)) || launch_ring_info(c, (int)nd, dlist, dlen, dbase) ||
        download(c, len.data(), dlen, nd * sizeof(uint32_t)))
        return -1;
    std::vector<int32_t> o(nd + 1, 0);
    for (uint32_t i = 0; i < nd; i++) {
        if (len[i] == 0) {
            c->err = "lbib_perturbation_rings: a flagged connection's domain has no ring";
            return -1;
        }
        o[i + 1] = o[i] + (int32_t)len[i];
        if ((long)o[i + 1] > (long)INT_MAX / 2) { c->err = "lbib_perturbation_rings: rings too long"; return -1; }
    }
    *nnodes = o[nd];
    for (uint32_t i = 0; i < nd; i++) domain[i] = doms[i];
    for (uint32_t i = 0; i <= nd; i++) off[i] = o[i];
    if (o[nd] > cap_nodes) return 0;
    if (upload(c, doff, o.data(), (nd + 1) * sizeof(int32_t)) ||
        launch_ring_gather(c, (int)nd, dbase, doff, dmarker, dxy) ||
        download(c, marker, dmarker, (size_t)o[nd] * sizeof(int32_t)) ||
        download(c, xy, dxy, 2 * (size_t)o[nd] * sizeof(double)))
        return -1;
    return 0;
}
int lbib_set_marker_positions(lbib_ctx* c, int n, const int32_t* index, const double* xy) {
    if (check_ctx(c) || n < 0 || (n > 0 && (!index || !xy))) return -1;
    if (n == 0) return 0;
    if (sync_counts(c)) return -1;
    for (int i = 0; i < n; i++)
        if (index[i] < 0 || index[i] >= c->nm) {
            c->err = "lbib_set_marker_positions: marker index outside the live markers";
            return -1;
        }
    if (check_marker_edges(c, n, xy, index)) return -1;
    Carve sz{nullptr};
    sz.take<double>(2 * (size_t)n); sz.take<int32_t>((size_t)n);
    Carve cv{pert_staging(c, sz.at)};
    if (!cv.base) return -1;
    double* dxy = cv.take<double>(2 * (size_t)n);
    int32_t* dix = cv.take<int32_t>((size_t)n);
    if (upload(c, dxy, xy, 2 * (size_t)n * sizeof(double)) || upload(c, dix, index, (size_t)n * sizeof(int32_t)) ||
        launch_set_positions(c, n, dix, dxy))
        return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
int lbib_set_marker_ids(lbib_ctx* c, const uint32_t* ids) {
    if (check_ctx(c) || (c->nm > 0 && !ids)) return -1;
    uint32_t mx = 0;
    for (int k = 0; k < c->nm; k++) {
        if (k > 0 && ids[k] <= ids[k - 1]) { c->err = "marker ids must increase with the dense index"; return -1; }
        mx = std::max(mx, ids[k]);
    }
    if (upload(c, c->mid, ids, c->nm * sizeof(uint32_t))) return -1;
    LBIB_HIP(hipMemcpyAsync((char*)c->st + offsetof(DevStatus, maxid), &mx, sizeof(uint32_t), hipMemcpyHostToDevice,
                            c->stream));
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
int lbib_get_marker_ids(lbib_ctx* c, uint32_t* ids) {
    if (check_ctx(c) || (c->nm > 0 && !ids)) return -1;
    return download(c, ids, c->mid, c->nm * sizeof(uint32_t));
}
int lbib_get_connections(lbib_ctx* c, int32_t* a, int32_t* b, uint32_t* dom, int32_t* bc) {
    if (check_ctx(c)) return -1;
    const size_t n = (size_t)c->nc;
    if ((a && download(c, a, c->ca, n * sizeof(int32_t))) || (b && download(c, b, c->cb, n * sizeof(int32_t))) ||
        (dom && download(c, dom, c->cdom, n * sizeof(uint32_t))) || (bc && download(c, bc, c->cbc, n * sizeof(int32_t))))
        return -1;
    return 0;
}
int lbib_get_forces(lbib_ctx* c, int32_t* off, int32_t* type, int32_t* partner, double* prm) {
    if (check_ctx(c) || !off) return -1;
    if (download(c, off, c->fs_off, (c->nm + 1) * sizeof(int32_t))) return -1;
    const int nf = off[c->nm];
    if (nf > c->fcap) { c->err = "force table larger than its capacity"; return -1; }
    if ((type && download(c, type, c->fs_type, nf * sizeof(int32_t))) ||
        (partner && download(c, partner, c->fs_partner, nf * sizeof(int32_t))) ||
        (prm && download(c, prm, c->fs_prm, 4 * (size_t)nf * sizeof(double))))
        return -1;
    return 0;
}
int lbib_get_markers(lbib_ctx* c, double* xy, double* vxy, double* fxy, double* oxy) {
    if (check_ctx(c)) return -1;
    const int nm = c->nm;
    std::vector<double> a(nm), b(nm);
    auto join = [&](double* dst, const double* dx, const double* dy) -> int {
        if (download(c, a.data(), dx, nm * sizeof(double)) || download(c, b.data(), dy, nm * sizeof(double))) return -1;
        for (int k = 0; k < nm; k++) { dst[2 * k] = a[k]; dst[2 * k + 1] = b[k]; }
        return 0;
    };
    if (nm == 0) return 0;
    if (xy && join(xy, c->mx, c->my)) return -1;
    if (vxy && join(vxy, c->mvx, c->mvy)) return -1;
    if (fxy && join(fxy, c->mfx, c->mfy)) return -1;
    if (oxy && join(oxy, c->mox, c->moy)) return -1;
    return 0;
}
int lbib_set_connections(lbib_ctx* c, int nc, const int32_t* a, const int32_t* b, const uint32_t* dom,
                         const int32_t* bc) {
    if (check_ctx(c) || nc < 0) return -1;
    if (nc > 0 && (!a || !b || !dom)) return -1;
    for (int k = 0; k < nc; k++) {
        if (a[k] < 0 || a[k] >= c->nm || b[k] < 0 || b[k] >= c->nm) {
            c->err = "connection refers to a marker index out of range";
            return -1;
        }
        if (dom[k] > DOM_MASK) { c->err = "domain id exceeds 24 bits"; return -1; }
        if (bc && bc[k] >= c->nbcd) { c->err = "connection refers to an unknown BC descriptor"; return -1; }
    }
    if (nc > c->ccap && grow_connections(c, nc)) return -1;
    if (upload(c, c->ca, a, nc * sizeof(int32_t)) || upload(c, c->cb, b, nc * sizeof(int32_t)) ||
        upload(c, c->cdom, dom, nc * sizeof(uint32_t)))
        return -1;
    std::vector<int32_t> none;
    const int32_t* bcp = bc;
    if (!bcp) { none.assign(std::max(nc, 1), -1); bcp = none.data(); }
    if (upload(c, c->cbc, bcp, nc * sizeof(int32_t))) return -1;
    c->nc = nc;
    uint32_t maxdom = 0;
    for (int k = 0; k < nc; k++) maxdom = std::max(maxdom, dom[k]);
    if ((int)maxdom + 1 > c->ndom) {
        if (ensure_domains(c, (int)maxdom + 1)) return -1;
        c->ndom = (int)maxdom + 1;
    }
    if (put_counts(c)) return -1;
    // marker in/out links and ring order of the new topology
    if (launch_inout(c) || launch_rings(c)) return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
int lbib_set_bc_descriptors(lbib_ctx* c, int nd, const int32_t* off, const int32_t* species) {
    if (check_ctx(c) || nd < 0 || !off) return -1;
    const int total = off[nd];
    for (int k = 0; k < total; k++)
        if (species[k] < 0 || species[k] >= c->ns) { c->err = "BC descriptor species out of range"; return -1; }
    dev_free(c->bcd_off);
    dev_free(c->bcd_sp);
    if (dev_alloc(c, &c->bcd_off, nd + 1) || dev_alloc(c, &c->bcd_sp, std::max(total, 1))) return -1;
    if (upload(c, c->bcd_off, off, (nd + 1) * sizeof(int32_t)) || upload(c, c->bcd_sp, species, total * sizeof(int32_t)))
        return -1;
    c->nbcd = nd;
    if (nd > 0 && !c->bc_v) {
        if (dev_alloc(c, &c->bc_v, 8 * c->n) || dev_alloc(c, &c->bc_desc, 4 * c->n) ||
            dev_alloc(c, &c->cand_v, (size_t)c->candcap))
            return -1;
    }
    c->any_bc = nd > 0;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
int lbib_set_celltypes(lbib_ctx* c, int ndom, const uint32_t* celltype) {
    if (check_ctx(c) || ndom < 0 || (ndom > 0 && !celltype)) return -1;
    if (ensure_domains(c, ndom)) return -1;
    if (upload(c, c->celltype, celltype, ndom * sizeof(uint32_t))) return -1;
    c->ndom = ndom;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
int lbib_set_forces(lbib_ctx* c, int nm, const int32_t* off, const int32_t* type, const int32_t* partner,
                    const double* prm) {
    if (check_ctx(c) || nm != c->nm || !off) { if (c) c->err = "force CSR must cover every marker"; return -1; }
    const int nf = off[nm];
    for (int j = 0; j < nf; j++) {
        if (type[j] < 0 || type[j] > 7) { c->err = "Invalid force solver type found."; return -1; }
        const bool pair = type[j] == 0 || type[j] == 5 || type[j] == 6 || type[j] == 7;
        if (pair && (partner[j] < 0 || partner[j] >= nm)) { c->err = "Could not compute and set the forces."; return -1; }
    }
    if (ensure_force_cap(c, nf)) return -1;
    if (upload(c, c->fs_off, off, (nm + 1) * sizeof(int32_t)) || upload(c, c->fs_type, type, nf * sizeof(int32_t)) ||
        upload(c, c->fs_partner, partner, nf * sizeof(int32_t)) || upload(c, c->fs_prm, prm, 4 * nf * sizeof(double)))
        return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ------------------------------------------------------------------- step --
static int next_event(lbib_ctx* c, hipEvent_t* ev) {
    if (c->ev_used == c->ev_pool.size()) {
        // device-scope release: the timing pair brackets one kernel on this
        // device without a system-scope cache write-back around it
        hipEvent_t e;
        LBIB_HIP(hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
        c->ev_pool.push_back(e);
    }
    *ev = c->ev_pool[c->ev_used++];
    return 0;
}

static int stream_collide(lbib_ctx* c, uint32_t it, bool ftile) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->capturing) it = ITER_NEXT;
    const bool timed = c->prof_every > 0 && !c->capturing && (c->prof_seq++ % (uint64_t)c->prof_every) == 0;
    if (timed) {
        if (next_event(c, &e0) || next_event(c, &e1)) return -1;
        LBIB_HIP(hipEventRecord(e0, c->stream));
    }
    const int r = c->prec == LBIB_FP64 ? launch_stream_collide_f64(c, it, ftile) : launch_stream_collide_mixed(c, it, ftile);
    if (r) return r;
    if (timed) LBIB_HIP(hipEventRecord(e1, c->stream));
    return 0;
}

int lbib_step_lattice(lbib_ctx* c, uint32_t it) {
    if (check_ctx_keep_begin(c)) return -1;   // the first lattice kernel opens the iteration
    if (launch_force_and_spread(c)) return -1;
    if (stream_collide(c, it, c->nm > 0)) return -1;
    if (launch_clear_force(c)) return -1;
    if (launch_interp_move(c)) return -1;
    return 0;
}
int lbib_step_lattice_only(lbib_ctx* c, uint32_t it) {
    if (check_ctx(c)) return -1;
    return stream_collide(c, it, false);
}
int lbib_step_geometry(lbib_ctx* c, uint32_t it, int coarsen) {
    if (check_ctx(c)) return -1;
    return launch_geometry(c, c->capturing ? ITER_NEXT : it, coarsen ? 2 : 0);
}
int lbib_step_topology(lbib_ctx* c, int coarsen) {
    if (check_ctx(c)) return -1;
    return launch_topology(c, coarsen != 0);
}
int lbib_forces_tension(lbib_ctx* c, double f) {
    if (check_ctx(c)) return -1;
    if (ensure_force_cap(c, 2 * c->mcap)) return -1;
    return launch_forces_pairs(c, 6, f, 0.0);
}
int lbib_forces_springs(lbib_ctx* c, double k, double l0) {
    if (check_ctx(c)) return -1;
    if (ensure_force_cap(c, 2 * c->mcap)) return -1;
    return launch_forces_pairs(c, 0, k, l0);
}
int lbib_mass_growth(lbib_ctx* c, int species_r, int species_l, double threshold, double maxsource) {
    if (check_ctx(c)) return -1;
    if (species_r < 0 || species_r >= c->ns || species_l < 0 || species_l >= c->ns) {
        c->err = "lbib_mass_growth: species index out of range";
        return -1;
    }
    return launch_mass_growth(c, species_r, species_l, threshold, maxsource);
}
int lbib_forces_junctions(lbib_ctx* c, double radius, double k, double l0) {
    if (check_ctx(c)) return -1;
    if (!std::isfinite(radius) || radius < 0.0 || !std::isfinite(k) || !std::isfinite(l0)) {
        c->err = "lbib_forces_junctions: radius must be finite and >= 0, k and l0 finite";
        return -1;
    }
    if (c->mcap == 0) return 0;
    c->junc_on = true;
    if (ensure_junctions(c)) return -1;
    return launch_forces_junctions(c, radius, k, l0);
}
int lbib_forces_clear(lbib_ctx* c) {
    if (check_ctx(c)) return -1;
    return launch_forces_clear(c);
}
int lbib_cure_lattice(lbib_ctx* c) {
    if (check_ctx(c)) return -1;
    return launch_geometry(c, 0, 1);
}
int lbib_profile_enable(lbib_ctx* c, int on) {
    if (check_ctx(c)) return -1;
    c->prof_every = on > 0 ? on : 0;
    c->prof_seq = 0;
    return 0;
}
int lbib_profile_read(lbib_ctx* c, double* kernel_ms, int64_t* launches, double* bytes_per_launch) {
    if (check_ctx(c)) return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    double ms = 0.0;
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        float t = 0.0f;
        LBIB_HIP(hipEventElapsedTime(&t, c->ev_pool[k], c->ev_pool[k + 1]));
        ms += t;
    }
    if (kernel_ms) *kernel_ms = ms;
    if (launches) *launches = (int64_t)(c->ev_used / 2);
    if (bytes_per_launch)
        *bytes_per_launch = 2.0 * (double)(field_bytes(c->prec, 9) + c->ns * field_bytes(c->prec, 5)) *
                            (double)c->live_sites;
    c->ev_used = 0;
    return 0;
}
int lbib_profile_copy(lbib_ctx* c, int64_t nbytes, int reps, double* kernel_ms) {
    if (check_ctx(c)) return -1;
    if (nbytes < 4096 || reps < 1 || !kernel_ms) {
        c->err = "lbib_profile_copy: nbytes >= 4096, reps >= 1 and an output pointer are required";
        return -1;
    }
    nbytes &= ~(int64_t)15;
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t e[2] = {nullptr, nullptr};
    int rc = -1;
    float t = 0.0f;
    if (hipMalloc(&buf[0], (size_t)nbytes) != hipSuccess || hipMalloc(&buf[1], (size_t)nbytes) != hipSuccess ||
        hipEventCreate(&e[0]) != hipSuccess || hipEventCreate(&e[1]) != hipSuccess) {
        c->err = "lbib_profile_copy: allocation failed";
        goto done;
    }
    if (hipMemsetAsync(buf[0], 0, (size_t)nbytes, c->stream) != hipSuccess) goto hip_fail;
    if (launch_copy_probe(c, buf[0], buf[1], (long)nbytes)) goto done;  // warm-up
    if (hipEventRecord(e[0], c->stream) != hipSuccess) goto hip_fail;
    for (int r = 0; r < reps; r++)
        if (launch_copy_probe(c, buf[r & 1], buf[1 - (r & 1)], (long)nbytes)) goto done;
    if (hipEventRecord(e[1], c->stream) != hipSuccess || hipEventSynchronize(e[1]) != hipSuccess ||
        hipEventElapsedTime(&t, e[0], e[1]) != hipSuccess)
        goto hip_fail;
    *kernel_ms = (double)t / reps;
    rc = 0;
    goto done;
hip_fail:
    c->err = std::string("lbib_profile_copy: ") + hipGetErrorString(hipGetLastError());
done:
    (void)hipStreamSynchronize(c->stream);
    for (auto* b : buf)
        if (b) (void)hipFree(b);
    for (auto ev : e)
        if (ev) (void)hipEventDestroy(ev);
    return rc;
}
int lbib_reserve(lbib_ctx* c, int max_markers, int max_connections) {
    if (check_ctx(c) || max_markers < c->nm || max_connections < c->nc) {
        if (c) c->err = "reserve below the current counts";
        return -1;
    }
    // reallocation drops the topology: the caller uploads it again
    if (ensure_markers(c, max_markers)) return -1;
    if (max_connections > c->ccap && grow_connections(c, max_connections)) return -1;
    return 0;
}
int lbib_set_box_outlet(lbib_ctx* c, int on) {
    if (check_ctx(c)) return -1;
    c->cfg.box_outlet = on ? 1 : 0;
    return 0;
}
int lbib_read_status(lbib_ctx* c, lbib_status* out) {
    if (check_ctx(c) || !out) return -1;
    const size_t nsh = (size_t)SH_COUNT * NSHARD * SHARD_STRIDE;
    LBIB_HIP(hipMemcpyAsync(c->st_host, c->st, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream));
    LBIB_HIP(hipMemcpyAsync(c->shard_host, c->shard, nsh * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    LBIB_HIP(hipStreamSynchronize(c->stream));
    const DevStatus& s = *c->st_host;
    out->bits = (s.bits & ERR_EDGE_REMESH) ? ((s.bits & ~ERR_EDGE_REMESH) | LBIB_ERR_EDGE) : s.bits;
    out->first_bad_marker = s.first_bad == INT_MAX ? -1 : s.first_bad;
    uint32_t touched = 0;
    for (int k = 0; k < NSHARD; k++) touched += c->shard_host[(SH_TOUCHED * NSHARD + k) * SHARD_STRIDE];
    out->n_touched_sites = (int32_t)touched;
    uint32_t nbn = 0;
    for (int k = 0; k < NSHARD; k++) nbn += c->shard_host[(SH_NBN * NSHARD + k) * SHARD_STRIDE];
    out->n_boundary_nodes = (int32_t)nbn;
    out->n_reinit = (int32_t)s.nchanged;
    out->n_candidates = (int32_t)s.ncand;
    out->n_markers = (int32_t)s.nm;
    out->n_connections = (int32_t)s.nc;
    out->n_remeshed = (int32_t)s.rm_total;
    out->n_coarsened = (int32_t)s.co_removed;
    out->step_iteration = s.step_it;
    out->n_reinit_levels = (int32_t)s.reinit_levels;
    c->nm = (int)s.nm;
    c->nc = (int)s.nc;
    return 0;
}
int lbib_step_begin(lbib_ctx* c, uint32_t it) {
    if (check_ctx_keep_forces(c)) return -1;
    c->begin_pending = true;
    c->begin_it = it;
    if (!c->snaps.empty() && c->snaps.back().it >= it) c->snaps.clear();   // a new run
    c->snaps.push_back({it, c->cur, c->mtab_live, c->msite_live, c->ev_used});
    if (c->capturing) c->gsnaps.push_back(c->snaps.back());
    if (c->snaps.size() > 1024) c->snaps.pop_front();
    return 0;
}
int lbib_rewind(lbib_ctx* c, uint32_t it) {
    if (check_ctx_keep_begin(c)) return -1;
    if (c->begin_pending && c->begin_it >= it) c->begin_pending = false;   // never reached the device
    c->forces_beside = false;
    if (join_topology(c)) return -1;
    if (flush_step_begin(c)) return -1;
    if (c->snaps.empty() || it > c->snaps.back().it) return 0;   // iteration it never began
    while (!c->snaps.empty() && c->snaps.back().it > it) c->snaps.pop_back();
    if (c->snaps.empty() || c->snaps.back().it != it) {
        c->snaps.clear();
        c->err = "lbib_rewind: no snapshot for iteration " + std::to_string(it);
        return -1;
    }
    const lbib_ctx::Snapshot sn = c->snaps.back();
    c->snaps.pop_back();
    c->cur = sn.cur;
    c->mtab_live = sn.mtab_live;
    c->msite_live = sn.msite_live;
    c->ev_used = sn.ev_used;
    return 0;
}
int lbib_clear_status(lbib_ctx* c) {
    if (check_ctx(c)) return -1;
    // bits and first_bad are the first two words of DevStatus
    struct { uint32_t bits; int32_t first; } z = {0u, INT_MAX};
    LBIB_HIP(hipMemcpyAsync(c->st, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// -------------------------------------------------------------- bio helpers --
int lbib_compute_areas(lbib_ctx* c, double* area) {
    if (check_ctx(c)) return -1;
    if (launch_ring_areas(c, nullptr)) return -1;
    if (area && c->ndom > 0) {
        std::vector<uint32_t> cnt(c->ndom);
        if (download(c, cnt.data(), c->ring_cnt, c->ndom * sizeof(uint32_t))) return -1;
        if (download(c, area, c->area, c->ndom * sizeof(double))) return -1;
        for (int d = 0; d < c->ndom; d++)
            if (!cnt[d]) area[d] = 0.0;
    }
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
int lbib_domain_concentrations(lbib_ctx* c, int species, int ndom, double* out) {
    if (check_ctx(c)) return -1;
    if (species < 0 || species >= c->ns || ndom < 1 || !out) {
        c->err = "lbib_domain_concentrations: bad species index or output size";
        return -1;
    }
    if (ndom > c->dconc_cap) {
        dev_free(c->dconc);
        c->dconc_cap = 0;
        if (dev_alloc(c, &c->dconc, ndom)) return -1;
        c->dconc_cap = ndom;
    }
    LBIB_HIP(hipMemsetAsync(c->dconc, 0, ndom * sizeof(double), c->stream));
    if (launch_domain_concentrations(c, species, ndom, c->dconc)) return -1;
    return download(c, out, c->dconc, ndom * sizeof(double));
}
int lbib_area_regulator(lbib_ctx* c, double target_area, double kp, double maxsource, double target_factor) {
    if (check_ctx_keep_forces(c)) return -1;
    return launch_area_regulator(c, target_area, kp, maxsource, target_factor);
}
int lbib_add_mass_by_domain(lbib_ctx* c, int ndom, const double* table) {
    if (check_ctx(c) || ndom < 0 || !table) return -1;
    if (ensure_domains(c, ndom)) return -1;
    if (ndom > c->ndom) c->ndom = ndom;
    std::vector<double> t(c->ndom, 0.0);
    for (int d = 0; d < ndom; d++) t[d] = table[d];
    if (upload(c, c->mtab, t.data(), c->ndom * sizeof(double))) return -1;
    c->mtab_live = true;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int lbib_add_mass_sites(lbib_ctx* c, int64_t count, const int64_t* site, const double* m) {
    if (check_ctx(c) || count < 0 || (count > 0 && (!site || !m))) return -1;
    // a site listed twice keeps its last value (addMass overwrites)
    std::vector<int64_t> s;
    std::vector<double> v;
    s.reserve((size_t)count);
    v.reserve((size_t)count);
    {
        for (int64_t k = 0; k < count; k++) {
            if (site[k] < 0 || site[k] >= c->n) {
                c->err = "lbib_add_mass_sites: site index outside the lattice";
                return -1;
            }
        }
        std::unordered_map<int64_t, size_t> at;
        at.reserve((size_t)count * 2);
        for (int64_t k = 0; k < count; k++) {
            auto it = at.find(site[k]);
            if (it == at.end()) {
                at.emplace(site[k], s.size());
                s.push_back(site[k]);
                v.push_back(m[k]);
            } else {
                v[it->second] = m[k];
            }
        }
    }
    const long cnt = (long)s.size();
    if (cnt > c->mass_cap) {
        dev_free(c->mass_site);
        dev_free(c->mass_val);
        c->mass_cap = 0;
        if (dev_alloc(c, &c->mass_site, (size_t)cnt) || dev_alloc(c, &c->mass_val, (size_t)cnt)) return -1;
        c->mass_cap = cnt;
    }
    if (upload(c, c->mass_site, s.data(), cnt * sizeof(int64_t)) || upload(c, c->mass_val, v.data(), cnt * sizeof(double)))
        return -1;
    if (launch_mass_sites(c, cnt, c->mass_site, c->mass_val)) return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ------------------------------------------------------------ batched atlas --
int lbib_set_batch_layout(lbib_ctx* c, int nrect, const lbib_rect* rects) {
    if (check_ctx(c)) return -1;
    if (nrect < 0 || (nrect > 0 && !rects)) { c->err = "lbib_set_batch_layout: bad arguments"; return -1; }
    const long nt = (long)c->stx * c->sty;
    std::vector<int32_t> owner(nt, -1);
    std::vector<RectDev> rd(nrect);
    long live = 0;
    bool pairs = true;
    for (int r = 0; r < nrect; r++) {
        const lbib_rect& q = rects[r];
        const std::string tag = "lbib_set_batch_layout: rectangle " + std::to_string(r) + ": ";
        if (q.x0 < 0 || q.y0 < 0 || q.x0 % 16 || q.y0 % 16 || q.w < 4 || q.h < 4 || q.x0 + q.w > c->nx ||
            q.y0 + q.h > c->ny) {
            c->err = tag + "origin must be a multiple of 16, sides >= 4, inside the lattice";
            return -1;
        }
        if (!(q.tau_fluid > 0.5)) { c->err = tag + "tau_fluid must exceed 0.5"; return -1; }
        for (int s = 0; s < c->ns; s++)
            if (!(q.tau_species[s] > 0.0)) { c->err = tag + "species tau must be positive"; return -1; }
        for (int ty = q.y0 / 16; ty < (q.y0 + q.h + 15) / 16; ty++)
            for (int tx = q.x0 / 16; tx < (q.x0 + q.w + 15) / 16; tx++) {
                int32_t& o = owner[(long)ty * c->stx + tx];
                if (o >= 0) {
                    c->err = tag + "shares a 16x16 tile with rectangle " + std::to_string(o);
                    return -1;
                }
                o = r;
            }
        RectDev& d = rd[r];
        std::memset(&d, 0, sizeof(d));
        d.x0 = q.x0; d.y0 = q.y0; d.w = q.w; d.h = q.h;
        d.tau = q.tau_fluid;
        d.tauI = 1.0 / q.tau_fluid;
        for (int s = 0; s < LBIB_MAX_SPECIES; s++) d.tauIs[s] = s < c->ns ? 1.0 / q.tau_species[s] : 1.0;
        live += (long)q.w * q.h;
        if (q.w % 2) pairs = false;
    }
    LBIB_HIP(hipStreamSynchronize(c->stream));
    dev_free(c->rtile);
    dev_free(c->rects);
    dev_free(c->vsegs);
    c->vseg_len = 0;
    c->nvseg = 0;
    c->nrect = 0;
    c->live_sites = c->n;
    c->rect_pairs = true;
    c->rect_host.clear();
    if (nrect == 0) return 0;
    for (int r = 0; r < nrect; r++) {
        const lbib_rect& q = rects[r];
        c->rect_host.insert(c->rect_host.end(), {q.x0, q.y0, q.w, q.h});
    }
    if (dev_alloc(c, &c->rtile, nt) || dev_alloc(c, &c->rects, nrect)) return -1;
    if (upload(c, c->rtile, owner.data(), nt * sizeof(int32_t)) ||
        upload(c, c->rects, rd.data(), nrect * sizeof(RectDev)))
        return -1;
    LBIB_HIP(hipStreamSynchronize(c->stream));
    c->nrect = nrect;
    c->live_sites = live;
    c->rect_pairs = pairs;
    return 0;
}

// ---------------------------------------------------------- launch shapes --
static bool shape_valid(const lbib_ctx* c, const LatticeShape& s) {
    const bool wg_ok = s.wg == 64 || s.wg == 128 || s.wg == 256;
    if (s.variant == 1) return wg_ok && pairs_apply(c);
    if (s.variant == 2) return wg_ok && (c->prec == LBIB_FP16_REM ? s.spt == 1 : vec_apply(c, s.spt));
    if (s.variant != 0 || !wg_ok) return false;
    return (c->prec == LBIB_FP64 || c->prec == LBIB_FP16_REM) ? s.spt == 1 : (s.spt == 1 || s.spt == 2 || s.spt == 4);
}
int lbib_set_lattice_shape(lbib_ctx* c, const lbib_lattice_shape* shape) {
    if (check_ctx(c) || !shape) return -1;
    const LatticeShape s{shape->variant, shape->workgroup, shape->sites_per_lane};
    if (!shape_valid(c, s)) {
        c->err = "lbib_set_lattice_shape: shape does not apply to this context";
        return -1;
    }
    c->lat_variant = s.variant; c->lat_wg = s.wg; c->lat_spt = s.spt;
    return 0;
}
int lbib_autotune(lbib_ctx* c, int reps, lbib_lattice_shape* tried, int max_tried, int* n_tried,
                  lbib_lattice_shape* chosen) {
    if (check_ctx(c)) return -1;
    if (reps < 1) reps = 1;
    std::vector<TunedShape> out;
    if ((c->prec == LBIB_FP64 ? autotune_f64(c, reps, out) : autotune_mixed(c, reps, out)) != 0) return -1;
    if (out.empty()) { c->err = "lbib_autotune: no shape applies"; return -1; }
    size_t best = 0;
    for (size_t k = 1; k < out.size(); k++)
        if (out[k].ms < out[best].ms) best = k;
    c->lat_variant = out[best].shape.variant;
    c->lat_wg = out[best].shape.wg;
    c->lat_spt = out[best].shape.spt;
    auto put = [](lbib_lattice_shape* d, const TunedShape& t) {
        d->variant = t.shape.variant; d->workgroup = t.shape.wg; d->sites_per_lane = t.shape.spt; d->ms = t.ms;
    };
    if (tried)
        for (int k = 0; k < (int)out.size() && k < max_tried; k++) put(&tried[k], out[k]);
    if (n_tried) *n_tried = (int)out.size();
    if (chosen) put(chosen, out[best]);
    return 0;
}

// ------------------------------------------------------------- step graphs --

int lbib_capture_begin(lbib_ctx* c) {
    if (check_ctx(c)) return -1;
    if (c->capturing) { c->err = "lbib_capture_begin: already recording"; return -1; }
    graph_drop(c);
    c->gkey = graph_key(c);
    LBIB_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    return 0;
}
int lbib_capture_end(lbib_ctx* c, int* ok) {
    if (!c) return -1;
    if (ok) *ok = 0;
    if (!c->capturing) { c->err = "lbib_capture_end: not recording"; return -1; }
    // the second stream rejoins the recorded stream; a step begin still
    // pending belongs to the recorded block
    c->forces_beside = false;
    const int fl = (join_topology(c) != 0 || flush_step_begin(c) != 0) ? -1 : 0;
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (fl != 0 || e != hipSuccess || g == nullptr) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        graph_drop(c);
        return 0;   // nothing was recorded that can be replayed; *ok = 0
    }
    c->graph = g;
    if (hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        graph_drop(c);
        return 0;
    }
    if (ok) *ok = 1;
    return 0;
}
int lbib_graph_launch(lbib_ctx* c, uint32_t it_first, int* stale) {
    if (check_ctx(c)) return -1;
    if (stale) *stale = 0;
    if (!c->gexec || c->capturing || graph_key(c) != c->gkey || c->gsnaps.empty()) {
        if (stale) *stale = 1;
        return 0;
    }
    // the launch state of every replayed step, for lbib_rewind
    const uint32_t it0 = c->gsnaps.front().it;
    for (const lbib_ctx::Snapshot& s : c->gsnaps) {
        lbib_ctx::Snapshot t = s;
        t.it = it_first + (s.it - it0);
        if (!c->snaps.empty() && c->snaps.back().it >= t.it) c->snaps.clear();
        c->snaps.push_back(t);
        if (c->snaps.size() > 1024) c->snaps.pop_front();
    }
    LBIB_HIP(hipGraphLaunch(c->gexec, c->stream));
    return 0;
}
int lbib_graph_reset(lbib_ctx* c) {
    if (!c) return -1;
    graph_drop(c);
    return 0;
}

}  // extern "C"
