// lbib_compat.cpp -- the reference's plugin signatures over the device path
// (declarations and citations in lbib_compat.hpp).
#include "lbib_compat.hpp"

#include <iostream>

namespace LbmLib {

GlobalSimulationParameters Parameters;

// ------------------------------------------------------------ BioSolvers ----
namespace solver {
namespace {
struct FactoryEntry {
    std::function<std::string()> name;
    BioSolverFactory::Creator create;
};
std::vector<FactoryEntry>& factoryEntries() {
    static std::vector<FactoryEntry> e;
    return e;
}
}  // namespace

bool BioSolverFactory::add(std::function<std::string()> name, Creator create) {
    factoryEntries().push_back({std::move(name), std::move(create)});
    return true;
}
std::unique_ptr<BioAbstractSolver> BioSolverFactory::create(const std::string& name) {
    for (auto& e : factoryEntries())
        if (e.name() == name) return e.create();
    return nullptr;
}

void FluidSolver::addMass(double m) { v_->addMass(i_, m); }
Field<double> FluidSolver::getVelocity() const { return v_->velocity(i_); }
double FluidSolver::getRho() const { return v_->rho(i_); }
double FluidSolver::getTau() const { return v_->fluidTau(); }

double CDEAbstractSolver::getC() const { return v_->conc(species_, i_); }
std::string CDEAbstractSolver::getName() const { return v_->speciesName(species_); }
double CDEAbstractSolver::getTau() const { return v_->speciesTau(species_); }
}  // namespace solver

// --------------------------------------------------------- physical nodes ---
namespace nodes {
unsigned PhysicalNode::getDomainIdentifier() const { return v_->domain(i_); }
unsigned PhysicalNode::getCellType() const { return v_->cellType(i_); }
int PhysicalNode::getXPos() const { return (int)(i_ % v_->nx()); }
int PhysicalNode::getYPos() const { return (int)(i_ / v_->nx()); }
solver::CDEAbstractSolver& PhysicalNode::getCDESolverSlow(const std::string& name) const {
    const int k = v_->species(name);
    if (k < 0) throw Exception("The CDE Solver with the name " + name + " does not exists.");
    return v_->cde(k, i_);
}
}  // namespace nodes

// ------------------------------------------------------------- the view -----
namespace detail {

SimulationRunner& LatticeView::runner() const { return gh_->runner(); }

void LatticeView::shape() {
    if (nx_ > 0) return;
    // the grid of GeometryHandler::generatePhysicalGrid (GeometryHandler.cpp:463-474)
    const GlobalSimulationParameters& p = Parameters;
    nx_ = (int)p.sizeX;
    ny_ = (int)p.sizeY;
    if (nx_ <= 0 || ny_ <= 0) throw Exception("the lattice size is not set (Parameters SizeX, SizeY)");
}

const std::vector<std::vector<nodes::PhysicalNode*>>& LatticeView::grid() {
    std::lock_guard<std::mutex> lk(mu_);
    if (!grid_.empty()) return grid_;
    shape();
    const long n = (long)nx_ * ny_;
    nodes_.resize((size_t)n);
    for (long i = 0; i < n; i++) {
        nodes::PhysicalNode& p = nodes_[(size_t)i];
        p.v_ = this;
        p.i_ = i;
        p.fluid_.v_ = this;
        p.fluid_.i_ = i;
    }
    grid_.assign((size_t)ny_, std::vector<nodes::PhysicalNode*>((size_t)nx_));
    for (int y = 0; y < ny_; y++)
        for (int x = 0; x < nx_; x++) grid_[(size_t)y][(size_t)x] = &nodes_[(size_t)y * nx_ + x];
    return grid_;
}

// copy one array from the device when the runner's state changed since it
// was last copied (waits for the enqueued iterations first)
template <typename F>
void LatticeView::refresh(std::atomic<uint64_t>& gen, F&& copy) {
    SimulationRunner& r = runner();
    const uint64_t now = r.stateGeneration();
    if (gen.load(std::memory_order_acquire) == now) return;
    std::lock_guard<std::mutex> lk(mu_);
    if (gen.load(std::memory_order_relaxed) == now) return;
    shape();
    r.settleDevice();
    copy(r.context(), (size_t)nx_ * ny_);
    gen.store(now, std::memory_order_release);
}

static void ok(SimulationRunner& r, int rc, const char* what) {
    if (rc != 0) throw Exception(std::string(what) + ": " + lbib_last_error(r.context()));
}

unsigned LatticeView::domain(long i) {
    refresh(genDom_, [&](lbib_ctx* c, size_t n) {
        dom_.resize(n);
        ok(runner(), lbib_get_domains(c, dom_.data()), "getDomainIdentifier");
    });
    return dom_[(size_t)i];
}

unsigned LatticeView::cellType(long i) {
    // copyCellTypeToPhysicalNodes runs every step (GeometryHandler.cpp:382):
    // the site's type is its domain's entry (1 for a domain not listed)
    const unsigned d = domain(i);
    const auto& m = gh_->getCellTypeTrackerMap();
    auto it = m.find(d);
    return it == m.end() ? 1u : it->second;
}

Field<double> LatticeView::velocity(long i) {
    refresh(genVel_, [&](lbib_ctx* c, size_t n) {
        ux_.resize(n);
        uy_.resize(n);
        ok(runner(), lbib_get_velocity(c, ux_.data(), uy_.data()), "getVelocity");
    });
    return Field<double>(ux_[(size_t)i], uy_[(size_t)i]);
}

// the stored state keeps direction INV[q] in slot q (FluidSolver.cpp:259-264);
// getRho / getC sum the slots in index order
static const int kNatOfSlot9[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const int kNatOfSlot5[5] = {0, 3, 4, 1, 2};

double LatticeView::rho(long i) {
    refresh(genRho_, [&](lbib_ctx* c, size_t n) {
        std::vector<double> f(9 * n);
        ok(runner(), lbib_get_populations(c, f.data()), "getRho");
        rho_.assign(n, 0.0);
        for (size_t s = 0; s < n; s++) {
            double r = 0.0;
            for (int q = 0; q < 9; q++) r += f[(size_t)kNatOfSlot9[q] * n + s];
            rho_[s] = r;
        }
    });
    return rho_[(size_t)i];
}

double LatticeView::conc(int k, long i) {
    if (k < 0 || k >= LBIB_MAX_SPECIES) throw Exception("species index out of range");
    refresh(genC_[k], [&](lbib_ctx* c, size_t n) {
        std::vector<double> g(5 * n);
        ok(runner(), lbib_get_species(c, k, g.data()), "getC");
        c_[(size_t)k].assign(n, 0.0);
        for (size_t s = 0; s < n; s++) {
            double C = 0.0;
            for (int q = 0; q < 5; q++) C += g[(size_t)kNatOfSlot5[q] * n + s];
            c_[(size_t)k][s] = C;
        }
    });
    return c_[(size_t)k][(size_t)i];
}

int LatticeView::species(const std::string& name) const { return runner().speciesIndex(name); }
const std::string& LatticeView::speciesName(int k) const { return runner().speciesNames().at((size_t)k); }
double LatticeView::speciesTau(int k) const {
    return runner().parameters().cdeSolvers.at(runner().speciesNames().at((size_t)k));
}
double LatticeView::fluidTau() const { return runner().parameters().tauFluid; }

solver::CDEAbstractSolver& LatticeView::cde(int k, long i) {
    std::lock_guard<std::mutex> lk(mu_);
    if ((int)cde_.size() <= k) cde_.resize((size_t)k + 1);
    auto& v = cde_[(size_t)k];
    if (v.empty()) {
        shape();
        v.resize((size_t)nx_ * ny_);
        for (size_t s = 0; s < v.size(); s++) {
            v[s].v_ = this;
            v[s].i_ = (long)s;
            v[s].species_ = k;
        }
    }
    return v[(size_t)i];
}

void LatticeView::addMass(long i, double m) {
    std::lock_guard<std::mutex> lk(mu_);
    massSite_.push_back(i);
    massVal_.push_back(m);
}

void LatticeView::flushMass() {
    std::vector<int64_t> s;
    std::vector<double> v;
    {
        std::lock_guard<std::mutex> lk(mu_);
        s.swap(massSite_);
        v.swap(massVal_);
    }
    if (s.empty()) return;
    SimulationRunner& r = runner();
    ok(r, lbib_add_mass_sites(r.context(), (int64_t)s.size(), s.data(), v.data()), "addMass");
}

void LatticeView::discardMass() {
    std::lock_guard<std::mutex> lk(mu_);
    massSite_.clear();
    massVal_.clear();
}

}  // namespace detail

// ------------------------------------------------------- GeometryHandler ----
namespace geometry {

GeometryHandler::GeometryHandler(const Geometry& geometry)
    : geo_(const_cast<Geometry&>(geometry)), view_(new detail::LatticeView(this)) {}

GeometryHandler::~GeometryHandler() {
    if (runner_) runner_->detachGeometryHandler();
}

SimulationRunner& GeometryHandler::runner() const {
    if (!runner_) throw Exception("the GeometryHandler is not attached to a SimulationRunner");
    return *runner_;
}

const std::vector<std::vector<nodes::PhysicalNode*>>& GeometryHandler::getPhysicalNodes() const {
    return view_->grid();
}

const std::map<unsigned int, double> GeometryHandler::computeAreas() const { return runner().computeAreas(); }

std::map<unsigned int, double> GeometryHandler::computeAccumulatedDomainConcentrations(const std::string& name) const {
    return runner().computeAccumulatedDomainConcentrations(name);
}

void GeometryHandler::cureLattice() { runner().cureLattice(); }

void GeometryHandler::copyCellTypeToPhysicalNodes(std::map<unsigned int, unsigned int>& celltrackermap) {
    auto& own = geo_.cellTypeTrackerMap();
    if (&celltrackermap != &own) own = celltrackermap;
    runner().copyCellTypeToPhysicalNodes();
    // domains missing from the map were added as type 1
    if (&celltrackermap != &own) celltrackermap = own;
}

void GeometryHandler::copyCellTypeToPhysicalNodes(unsigned int domainidentifier, unsigned int celltype) {
    geo_.cellTypeTrackerMap()[domainidentifier] = celltype;
    runner().copyCellTypeToPhysicalNodes();
}

std::map<unsigned int, unsigned int>& GeometryHandler::getCellTypeTrackerMap(void) { return geo_.cellTypeTrackerMap(); }

void GeometryHandler::beginBioSolver() { view_->discardMass(); }
void GeometryHandler::endBioSolver() { view_->flushMass(); }

}  // namespace geometry

// --------------------------------------------------------------- reports ----
namespace reportHandler {
namespace {
SimulationRunner& runnerOf(const std::vector<std::vector<nodes::PhysicalNode*>>& nodes) {
    if (nodes.empty() || nodes[0].empty()) throw Exception("empty physical node grid");
    return nodes[0][0]->view()->runner();
}
}  // namespace

vtkFluidReporter::vtkFluidReporter(const std::vector<std::vector<nodes::PhysicalNode*>>& physicalNodes,
                                   const unsigned int cdecoarseningfactor, const std::string& filename)
    : AbstractReportFunctor(filename), physicalNodes_(physicalNodes), coarse_(cdecoarseningfactor) {}
void vtkFluidReporter::operator()(unsigned int time) const {
    runnerOf(physicalNodes_).writeVtkFluid(filename_, time, (int)coarse_);
}

vtkCDEReporter::vtkCDEReporter(const std::vector<std::vector<nodes::PhysicalNode*>>& physicalNodes,
                               const unsigned int cdecoarseningfactor, const std::string& filename)
    : AbstractReportFunctor(filename), physicalNodes_(physicalNodes), coarse_(cdecoarseningfactor) {}
void vtkCDEReporter::operator()(unsigned int time) const {
    runnerOf(physicalNodes_).writeVtkCDE(filename_, time, (int)coarse_);
}

vtkCellReporter::vtkCellReporter(const std::vector<std::shared_ptr<geometry::Connection>>& connections,
                                 const std::map<unsigned int, unsigned int>& celltypetrackermap,
                                 const std::string& filename)
    : AbstractReportFunctor(filename), connections_(connections), cellTypeTrackerMap_(celltypetrackermap) {}
void vtkCellReporter::operator()(unsigned int time) const {
    vtkxml::write_cells(filename_, time, connections_, cellTypeTrackerMap_);
}

}  // namespace reportHandler

// ------------------------------------------------------------ the runner ----
namespace {
// A BioSolver written against applyBioProcess(GeometryHandler&, ForceSolver&):
// it sees the global Parameters at the current iteration, and its addMass
// calls reach the device when it returns.
struct HandlerBioSolver : BioSolver {
    std::unique_ptr<solver::BioAbstractSolver> s;
    void applyBioProcess(SimulationRunner& r) override {
        geometry::GeometryHandler* gh = r.geometryHandler();
        Parameters.currentIteration = r.parameters().currentIteration;
        // a stop in the enqueued iterations is raised here, outside the
        // BioSolver (whose loops may be OpenMP regions)
        r.settleDevice();
        gh->beginBioSolver();
        s->applyBioProcess(*gh, r.getForceSolver());
        gh->endBioSolver();
    }
};
}  // namespace

SimulationRunner::SimulationRunner(const geometry::GeometryHandler& geometryHandler,
                                   const reportHandler::ReportHandler& reports, int precision)
    : SimulationRunner(Parameters, const_cast<geometry::Geometry&>(geometryHandler.getGeometry()), precision,
                       LBIB_FORCE_LUO) {
    gh_ = const_cast<geometry::GeometryHandler*>(&geometryHandler);
    report_ = &reports;
    gh_->runner_ = this;
    setup();
}

void SimulationRunner::addBioSolver(std::unique_ptr<solver::BioAbstractSolver> b) {
    if (!gh_) LBIB_FAIL("a BioAbstractSolver needs a SimulationRunner built on a GeometryHandler");
    std::unique_ptr<HandlerBioSolver> h(new HandlerBioSolver());
    h->name = b->getName();
    h->s = std::move(b);
    bio_.push_back(std::move(h));
}

void SimulationRunner::detachGeometryHandler() { gh_ = nullptr; }

}  // namespace LbmLib
