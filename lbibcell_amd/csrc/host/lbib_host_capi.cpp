// lbib_host_capi.cpp -- flat C entry points over the host runtime, used by the
// Python tests and bench (ctypes) and by any non-C++ caller.  They follow the
// reference's main() sequence (apps/tutorial_01.cpp:59-145): parameters,
// geometry, GeometryHandler (setup), SimulationRunner, initSolvers,
// initForceSolver, addMassSolver / addBioSolver, runSimulation.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <typeinfo>

#include "lbib_host.hpp"

using namespace LbmLib;

namespace {
struct Handle {
    GlobalSimulationParameters params;
    geometry::Geometry geo;
    std::unique_ptr<SimulationRunner> runner;
    std::map<std::string, lbib_species_spec> specs;
    std::vector<lbib_rect> rects;
    int precision = LBIB_FP64, forceModel = LBIB_FORCE_LUO;
    std::string err;
};

template <typename F>
int guarded(Handle* h, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::bad_typeid&) {
        h->err = "std::bad_typeid: unknown plugin name";
    } catch (const std::exception& e) {
        h->err = e.what();
    }
    return -1;
}

// the runner exists from lbh_setup on
SimulationRunner& R(Handle* h) {
    if (!h->runner) throw Exception("the simulation is not set up yet (call lbh_setup first)");
    return *h->runner;
}
}  // namespace

extern "C" {

void* lbh_create(int nx, int ny, double tauf, int force_model, int precision) {
    auto* h = new Handle();
    h->params.sizeX = (unsigned)nx;
    h->params.sizeY = (unsigned)ny;
    h->params.tauFluid = tauf;
    h->forceModel = force_model;
    h->precision = precision;
    return h;
}
void lbh_destroy(void* p) { delete static_cast<Handle*>(p); }
const char* lbh_last_error(void* p) { return static_cast<Handle*>(p)->err.c_str(); }

// Parameters "CDESolvers: name tau" plus an optional reaction spec for the name.
int lbh_add_species(void* p, const char* name, const lbib_species_spec* spec) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (h->runner) LBIB_FAIL("species must be added before setup");
        h->params.cdeSolvers[name] = spec->tau;
        // kind < 0: the registry's constants for that solver name
        if (spec->kind >= 0) h->specs[name] = *spec;
    });
}
int lbh_set_batch_layout(void* p, int n, const lbib_rect* rects) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (h->runner) LBIB_FAIL("the batch layout must be set before setup");
        // nodes draw their perturbations when added: from their domain's streams
        if (!h->geo.getGeometryNodes().empty()) LBIB_FAIL("the batch layout must be set before adding nodes");
        if (n < 0 || (n > 0 && !rects)) LBIB_FAIL("bad batch layout");
        h->rects.assign(rects, rects + n);
        h->geo.setBatchRects(h->rects);
    });
}
int lbh_load_parameters(void* p, const char* file) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { h->params.loadGlobalSimulationParameters(file); });
}
int lbh_load_geometry_vtk(void* p, const char* file) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { h->geo.loadGeometryVTK(file); });
}
// which: bit 0 fluid, bit 1 CDE solvers, bit 2 cells
int lbh_write_vtk(void* p, const char* prefix, unsigned time, int coarsening, int which) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (!h->runner) LBIB_FAIL("write_vtk needs a set-up runner");
        if (which & 1) R(h).writeVtkFluid(prefix, time, coarsening);
        if (which & 2) R(h).writeVtkCDE(prefix, time, coarsening);
        if (which & 4) R(h).writeVtkCells(prefix, time);
    });
}
// the host geometry as Geometry::writeGeometry text, or as a VTK cell block
// (vtkCellReporter layout): both work before setup
int lbh_write_geometry_txt(void* p, const char* file) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { h->geo.writeGeometry(file); });
}
int lbh_write_geometry_vtk(void* p, const char* prefix, unsigned time) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { vtkxml::write_cells(prefix, time, h->geo, h->geo.cellTypeTrackerMap()); });
}
int lbh_load_geometry_txt(void* p, const char* file) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { h->geo.loadGeometryTXT(file); });
}
int lbh_add_node(void* p, unsigned id, double x, double y) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { h->geo.addGeometryNode(id, x, y); });
}
// bc: comma-separated CDE names for a BoundarySolverNoFluxD2Q5 descriptor ("" = none)
int lbh_add_connection(void* p, unsigned id1, unsigned id2, unsigned dom, const char* bc) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        std::map<std::string, std::vector<std::string>> d;
        std::string s = bc ? bc : "";
        size_t pos = 0;
        while (!s.empty() && pos <= s.size()) {
            size_t q = s.find(',', pos);
            if (q == std::string::npos) q = s.size();
            if (q > pos) d["BoundarySolverNoFluxD2Q5"].push_back(s.substr(pos, q - pos));
            pos = q + 1;
        }
        h->geo.addConnection(id1, id2, dom, d);
    });
}
int lbh_set_celltype(void* p, unsigned dom, unsigned type) {
    auto* h = static_cast<Handle*>(p);
    h->geo.cellTypeTrackerMap()[dom] = type;
    return 0;
}
int lbh_setup(void* p) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        h->runner.reset(new SimulationRunner(h->params, h->geo, h->precision, h->forceModel));
        for (auto& kv : h->specs) R(h).setSpeciesSpec(kv.first, kv.second);
        R(h).setBatchLayout(h->rects);
        R(h).setup();
    });
}
int lbh_init_solvers(void* p, const double* u0) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).initSolvers(u0); });
}
int lbh_add_force(void* p, int type, unsigned id1, unsigned id2, double a, double b, double c, double d) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        solver::ForceStruct f{type, id1, id2, {a, b, c, d}};
        R(h).getForceSolver().addForce(f);
    });
}
int lbh_load_force_file(void* p, const char* file) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).initForceSolver(file); });
}
int lbh_add_mass_solver(void* p, const char* name) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).addMassSolver(name); });
}
int lbh_add_bio_solver(void* p, const char* name) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).addBioSolver(name); });
}
// Parameterised tutorial BioSolvers: kind 0 tension (prm[0] = force), kind 1
// area regulator (target, kp, maxsource, target factor).
int lbh_add_bio_solver_params(void* p, int kind, const double* prm, unsigned freq);
int lbh_write_simulation(void* p, const char* force, const char* geometry, const char* parameters,
                         const char* solvers) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).writeSimulation(force, geometry, parameters, solvers); });
}
int lbh_init_solvers_file(void* p, const char* solvers) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).initSolvers(std::string(solvers)); });
}
// the BioSolvers of the current iteration alone (SimulationRunner.cpp:95-97)
int lbh_apply_bio(void* p) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).applyBioSolvers(); });
}
int lbh_run(void* p, unsigned steps) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] { R(h).run(steps); });
}
// SimulationRunner::runSimulation (SimulationRunner.cpp:69-110) for `iterations`
// steps: a failure is caught, the reports and output/CRASH_*_<t>.txt files
// written (relative to the working directory) and the message printed
int lbh_run_simulation(void* p, unsigned iterations) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        R(h).parameters().iterations = iterations;
        R(h).runSimulation();
    });
}
int lbh_sync(void* p) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (lbib_synchronize(R(h).context()) != 0) throw Exception(lbib_last_error(R(h).context()));
    });
}
void* lbh_context(void* p) {
    auto* h = static_cast<Handle*>(p);
    return h->runner ? R(h).context() : nullptr;
}
unsigned lbh_iteration(void* p) {
    auto* h = static_cast<Handle*>(p);
    return h->runner ? R(h).parameters().currentIteration : h->params.currentIteration;
}
int lbh_set_iteration(void* p, unsigned it) {
    auto* h = static_cast<Handle*>(p);
    if (h->runner) R(h).parameters().currentIteration = it;
    else h->params.currentIteration = it;
    return 0;
}
static void mirror(Handle* h) {
    if (h->runner) R(h).syncMirror();
}
int lbh_num_markers(void* p) {
    auto* h = static_cast<Handle*>(p);
    if (guarded(h, [&] { mirror(h); })) return -1;
    return (int)h->geo.getGeometryNodes().size();
}
int lbh_num_connections(void* p) {
    auto* h = static_cast<Handle*>(p);
    if (guarded(h, [&] { mirror(h); })) return -1;
    return (int)h->geo.getConnections().size();
}
int lbh_get_markers(void* p, unsigned* ids, double* xy, double* v) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        mirror(h);
        size_t k = 0;
        for (auto& kv : h->geo.getGeometryNodes()) {
            if (ids) ids[k] = kv.first;
            if (xy) { xy[2 * k] = kv.second->x; xy[2 * k + 1] = kv.second->y; }
            if (v) { v[2 * k] = kv.second->vx; v[2 * k + 1] = kv.second->vy; }
            k++;
        }
    });
}
int lbh_get_connections(void* p, unsigned* a, unsigned* b, unsigned* dom) {
    auto* h = static_cast<Handle*>(p);
    if (guarded(h, [&] { mirror(h); })) return -1;
    size_t k = 0;
    for (auto& c : h->geo.getConnections()) {
        a[k] = c->first->id;
        b[k] = c->second->id;
        dom[k] = c->domain;
        k++;
    }
    return 0;
}
long lbh_perturb_counts(void* p, int which) {
    auto* h = static_cast<Handle*>(p);
    return which == 0 ? h->geo.nodePerturbations : h->geo.connectionPerturbations;
}
// host seconds (lattice, geometry, host fallbacks, bio enqueue; status
// waits), the step count, the number of resolved stops, and the node
// perturbations resolved from the flagged nodes / through the host geometry
void lbh_times(void* p, double* out) {
    auto* h = static_cast<Handle*>(p);
    if (!h->runner) {
        for (int k = 0; k < 9; k++) out[k] = 0.0;
        return;
    }
    const auto& t = R(h).times;
    out[0] = t.lattice; out[1] = t.geometry; out[2] = t.host; out[3] = t.bio; out[4] = (double)t.steps;
    out[5] = t.wait;
    out[6] = (double)t.stops;
    out[7] = (double)t.sparsePerturbations;
    out[8] = (double)t.fullPerturbations;
}
// iterations enqueued between status reads (1 = after every iteration)
int lbh_set_run_ahead(void* p, unsigned n) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (!h->runner) LBIB_FAIL("run-ahead is set after setup");
        R(h).runAhead = n < 1 ? 1 : n;
    });
}
// computeAccumulatedDomainConcentrations into out[0..cap) (domains absent: 0);
// returns the number of entries written (largest domain + 1)
int lbh_domain_concentrations(void* p, const char* name, double* out, int cap) {
    auto* h = static_cast<Handle*>(p);
    int written = -1;
    const int rc = guarded(h, [&] {
        auto m = R(h).computeAccumulatedDomainConcentrations(name);
        written = m.empty() ? 0 : (int)m.rbegin()->first + 1;
        for (int d = 0; d < cap; d++) out[d] = 0.0;
        for (auto& kv : m)
            if ((int)kv.first < cap) out[kv.first] = kv.second;
    });
    return rc != 0 ? -1 : written;
}
int lbh_set_graph_steps(void* p, unsigned n) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (!h->runner) LBIB_FAIL("graph steps are set after setup");
        R(h).setGraphSteps(n);
    });
}
// why graph mode turned itself off ("" while it is on or was never asked for)
int lbh_graph_note(void* p, char* buf, int n) {
    auto* r = static_cast<Handle*>(p)->runner.get();
    const std::string s = r ? r->graphNote : std::string();
    if (n > 0) {
        std::snprintf(buf, (size_t)n, "%s", s.c_str());
    }
    return (int)s.size();
}
// [steps per graph (0 once graph mode turned itself off), launches, recordings]
void lbh_graph_stats(void* p, long* out) {
    auto* r = static_cast<Handle*>(p)->runner.get();
    out[0] = r ? (long)r->graphSteps() : 0;
    out[1] = r ? r->graphLaunches : 0;
    out[2] = r ? r->graphCaptures : 0;
}
void lbh_reset_times(void* p) { static_cast<Handle*>(p)->runner->times = StageTimes{}; }
int lbh_last_status(void* p, lbib_status* st) {
    *st = static_cast<Handle*>(p)->runner->lastStatus;
    return 0;
}
int lbh_num_areas(void* p) {
    auto* h = static_cast<Handle*>(p);
    if (guarded(h, [&] { mirror(h); })) return -1;
    std::map<unsigned, int> d;
    for (auto& c : h->geo.getConnections()) d[c->domain] = 1;
    return (int)d.size();
}
int lbh_get_areas(void* p, unsigned* dom, double* area) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        auto a = R(h).computeAreas();
        size_t k = 0;
        for (auto& kv : a) { dom[k] = kv.first; area[k] = kv.second; k++; }
    });
}

}  // extern "C"

namespace {
struct TensionP : BioSolver {
    unsigned freq;
    double force;
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        r.renewTension(force);
    }
};
struct JunctionP : BioSolver {
    unsigned freq;
    double prm[3];
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        r.renewJunctions(prm[0], prm[1], prm[2]);
    }
};
struct AreaP : BioSolver {
    unsigned freq;
    double prm[4];
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        if (lbib_area_regulator(r.context(), prm[0], prm[1], prm[2], prm[3]) != 0)
            throw Exception(lbib_last_error(r.context()));
    }
};
}  // namespace

extern "C" int lbh_add_bio_solver_params(void* p, int kind, const double* prm, unsigned freq) {
    auto* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (freq == 0) LBIB_FAIL("frequency must be positive");
        if (kind == 0) {
            std::unique_ptr<TensionP> b(new TensionP());
            b->freq = freq;
            b->force = prm[0];
            b->name = "tutorial_01_BioSolverMembraneTension";
            R(h).addBioSolver(std::move(b));
        } else if (kind == 1) {
            std::unique_ptr<AreaP> b(new AreaP());
            b->freq = freq;
            for (int k = 0; k < 4; k++) b->prm[k] = prm[k];
            b->name = "tutorial_01_BioSolverAreaRegulator";
            R(h).addBioSolver(std::move(b));
        } else if (kind == 2) {
            if (!(prm[0] >= 0.0)) LBIB_FAIL("junction radius must be >= 0");
            std::unique_ptr<JunctionP> b(new JunctionP());
            b->freq = freq;
            for (int k = 0; k < 3; k++) b->prm[k] = prm[k];
            b->name = "tutorial_02_BioSolverCellJunction";
            R(h).addBioSolver(std::move(b));
        } else {
            R(h).addBioSolverWith(kind, prm, freq);
        }
    });
}
