// lbib_host.hpp -- host runtime of the device-resident lattice path, keeping
// the reference's plugin API: SimulationRunner, Geometry / GeometryHandler,
// ForceSolver, the CDE / Bio / Mass solver registries by name, and the global
// Parameters.  Membrane topology (node ids, connection order, remeshing, the
// seeded perturbation streams) is owned here; every per-site and per-marker
// computation runs on the GPU through the C ABI of include/lbib.h.
//
// Citations are relative to the reference repository root; "LbmLib/" means
// libs/LbmLib/.
#pragma once
#include <array>
#include <chrono>
#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/lbib.h"

namespace LbmLib {

// UtilLib::Exception / lbm_fail (UtilLib/include/Exception.hpp:90): message
// "file:line: msg".
struct Exception : std::runtime_error {
    explicit Exception(const std::string& m) : std::runtime_error(m) {}
};
#define LBIB_FAIL(msg) throw ::LbmLib::Exception(std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " + (msg))

// GlobalSimulationParameters (GlobalSimulationParameters.hpp:35-190).
struct GlobalSimulationParameters {
    unsigned currentIteration = 0;
    unsigned iterations = 0;
    unsigned reportSteps = 1;
    unsigned sizeX = 0, sizeY = 0;
    double tauFluid = 1.0;
    std::map<std::string, double> cdeSolvers;   // name -> tau (std::map: name order)
    void loadGlobalSimulationParameters(const std::string& file);   // :34-72
    void writeGlobalSimulationParameters(const std::string& file) const;
    // the reference's accessors (GlobalSimulationParameters.hpp:58-150)
    unsigned getCurrentIteration() const { return currentIteration; }
    unsigned getIterations() const { return iterations; }
    unsigned getReportSteps() const { return reportSteps; }
    void setReportSteps(unsigned r) { reportSteps = r; }
    unsigned getSizeX() const { return sizeX; }
    unsigned getSizeY() const { return sizeY; }
    double getTauFluid() const { return tauFluid; }
    const std::map<std::string, double>& getCdeSolvers() const { return cdeSolvers; }
    void printParameters() const;
    GlobalSimulationParameters& operator++() {
        ++currentIteration;
        return *this;
    }
    GlobalSimulationParameters operator++(int) {
        GlobalSimulationParameters old = *this;
        ++currentIteration;
        return old;
    }
};

namespace geometry {
class GeometryHandler;
}
namespace reportHandler {
class ReportHandler;
}
namespace solver {
class BioAbstractSolver;
}

namespace geometry {

struct Connection;
// GeometryNode (LbmLib/include/nodes/GeometryNode.hpp)
struct GeometryNode {
    unsigned id = 0;
    double x = 0, y = 0, vx = 0, vy = 0, ox = 0, oy = 0;
    Connection* in = nullptr;    // getConnection<0>
    Connection* out = nullptr;   // getConnection<1>
};

// Connection (LbmLib/include/geometry/Connection.hpp:47-155)
struct Connection {
    GeometryNode* first = nullptr;
    GeometryNode* second = nullptr;
    unsigned domain = 0;
    std::map<std::string, std::vector<std::string>> descriptor;   // boundary solver -> CDE names
};

// Geometry (LbmLib/include/geometry/Geometry.hpp:68-240): the membrane store.
class Geometry {
 public:
    Geometry() = default;
    explicit Geometry(const std::string& file);        // Geometry.cpp:56-72 (.txt, .vtm or .vtp)
    void loadGeometryTXT(const std::string& file);      // Geometry.cpp:99-160
    void loadGeometryVTK(const std::string& file);      // Geometry.cpp:162-306 (ASCII VTK XML)
    void writeGeometry(const std::string& file) const;  // Geometry.cpp:77-95
    // GeometryNode ctor + Geometry::addGeometryNode (Geometry.cpp:381-399)
    unsigned addGeometryNode(unsigned id, double x, double y);
    unsigned addGeometryNode(double x, double y);
    void addConnection(unsigned id1, unsigned id2, unsigned domain,
                       const std::map<std::string, std::vector<std::string>>& descriptor = {});
    void eraseConnection(Connection* c);                 // Geometry.cpp:471-491
    void removeGeometryNode(unsigned id);                // Geometry.cpp:401-428
    bool checkGeometryIntegrity() const;                 // Geometry.cpp:493-524
    // GeometryHandler::remeshBoundary / coarsenBoundary (GeometryHandler.cpp:385-461),
    // one pass each, same resulting node ids and connection order.
    unsigned remeshBoundary(double maxLength);
    unsigned coarsenBoundary(double maxLength);
    // replace nodes and connections (a device-side topology copied back); the
    // perturbation streams and cell types are kept
    void rebuild(const std::vector<unsigned>& ids, const std::vector<double>& xy, const std::vector<double>& vxy,
                 const std::vector<double>& oxy, const std::vector<int32_t>& a, const std::vector<int32_t>& b,
                 const std::vector<uint32_t>& dom,
                 const std::vector<std::map<std::string, std::vector<std::string>>>& descriptors);
    const std::map<unsigned, std::unique_ptr<GeometryNode>>& getGeometryNodes() const { return nodes_; }
    std::map<unsigned, std::unique_ptr<GeometryNode>>& nodes() { return nodes_; }
    const std::vector<std::shared_ptr<Connection>>& getConnections() const { return connections_; }
    std::map<unsigned, unsigned>& cellTypeTrackerMap() { return cellTypes_; }
    // seeded perturbation streams (GeometryNode.cpp:153-158, Connection.cpp:332-337)
    void perturbatePosition(GeometryNode& n);
    void setPos(GeometryNode& n, double x, double y);
    void perturbConnection(Connection& c);
    void perturbConnections();                           // GeometryHandler.cpp:326-356
    long nodePerturbations = 0, connectionPerturbations = 0;
    // Batched ensemble: each rectangle is an independent run, with perturbation
    // streams of its own seeded as a lone run's; a node draws from the
    // rectangle that contains it.
    void setBatchRects(const std::vector<lbib_rect>& rects);

 private:
    struct Streams {
        std::mt19937 node{1}, conn{1};
    };
    Streams& streamsAt(double x, double y);
    std::vector<lbib_rect> rects_;
    std::vector<Streams> rectStreams_;
    std::map<unsigned, std::unique_ptr<GeometryNode>> nodes_;
    std::vector<std::shared_ptr<Connection>> connections_;
    std::map<unsigned, unsigned> cellTypes_{{0u, 0u}};
    Streams global_;
    std::uniform_real_distribution<> disNode_{-10E-7, 10E-7};
    std::uniform_real_distribution<> disConn_{-10E-7, 10E-7};
};

}  // namespace geometry

namespace solver {
// ForceSolver (LbmLib/include/solver/ForceSolver.hpp:43-117)
struct ForceStruct {
    int type;
    unsigned id1, id2;
    double p[4];
};
class ForceSolver {
 public:
    void addForce(std::stringstream* descriptor);        // ForceSolver.cpp:107-139
    void addForce(const ForceStruct& f) { forcemap_[f.id1].push_back(f); dirty_ = true; }
    void loadForceFile(const std::string& file);         // ForceSolver.cpp:89-105
    void writeForceSolver(const std::string& file) const;
    void deleteAllForces() { forcemap_.clear(); dirty_ = true; }
    void replace(std::map<unsigned, std::vector<ForceStruct>>&& m) { forcemap_ = std::move(m); }
    bool onlyType(int t) const;
    void deleteForceType(unsigned type);                 // ForceSolver.cpp:168-204
    const std::map<unsigned, std::vector<ForceStruct>>& forcemap() const { return forcemap_; }
    bool dirty_ = true;

 private:
    std::map<unsigned, std::vector<ForceStruct>> forcemap_;
};
}  // namespace solver

class SimulationRunner;

// BioAbstractSolver::applyBioProcess (BioAbstractSolver.hpp:44-73)
struct BioSolver {
    virtual ~BioSolver() = default;
    virtual void applyBioProcess(SimulationRunner& runner) = 0;
    // A BioSolver whose applyBioProcess only enqueues device operations, and
    // decides on the iteration alone with this period, can be recorded into a
    // step graph (0: it cannot).
    virtual unsigned graphPeriod() const { return 0; }
    std::string name;
};

// Host seconds spent enqueueing each stage, waiting for the device at status
// reads, and in host fallbacks.
struct StageTimes {
    double lattice = 0, geometry = 0, host = 0, bio = 0, wait = 0;
    long steps = 0;
    long stops = 0;   // enqueued iterations that stopped and were resolved
    // perturbation halts resolved from the flagged nodes or rings alone, and
    // perturbation halts resolved through the whole host geometry
    long sparsePerturbations = 0, fullPerturbations = 0;
};

// SimulationRunner (LbmLib/include/SimulationRunner.hpp:43-231) fused with the
// GeometryHandler's device-facing duties.
class SimulationRunner {
 public:
    SimulationRunner(const GlobalSimulationParameters& params, geometry::Geometry& geo, int precision = LBIB_FP64,
                     int forceModel = LBIB_FORCE_LUO);
    // SimulationRunner(const GeometryHandler&, const ReportHandler&)
    // (SimulationRunner.hpp:51, SimulationRunner.cpp:50-64): the global
    // Parameters, the handler's geometry and its CDE solvers by name; the
    // lattice is set up at once (the GeometryHandler ctor's cureLattice,
    // GeometryHandler.cpp:51-56) and the handler serves its host view from it.
    // Defined in lbib_compat.cpp.
    SimulationRunner(const geometry::GeometryHandler& geometryHandler, const reportHandler::ReportHandler& reports,
                     int precision = LBIB_FP64);
    ~SimulationRunner();
    // CDE solver registry (CDESolverFactory): name -> reaction spec; called
    // before the context is created, in Parameters::cdeSolvers order.
    static bool knownCDESolver(const std::string& name);
    void setSpeciesSpec(const std::string& name, const lbib_species_spec& spec);
    // Batched ensemble: the lattice is an atlas of periodic rectangles, one
    // per independent domain (lbib_set_batch_layout); applied by setup().
    void setBatchLayout(const std::vector<lbib_rect>& rects) {
        rects_ = rects;
        geo_.setBatchRects(rects);
    }
    void setup();                                         // GeometryHandler ctor: cureLattice (:149-170)
    void initSolvers(const double* u0 = nullptr);         // SimulationRunner.cpp:290-320
    // SimulationRunner::initSolvers(const std::string&) (SimulationRunner.cpp:148-180):
    // the solver text dump written by writeSimulation
    void initSolvers(const std::string& solverFile);
    // SimulationRunner::writeSimulation (SimulationRunner.cpp:112-146): force,
    // geometry, parameter and solver text files in the reference's formats
    void writeSimulation(const std::string& forceFile, const std::string& geometryFile,
                         const std::string& parameterFile, const std::string& solverFile);
    void initForceSolver(const std::string& file);        // :392-394
    void addMassSolver(const std::string& name);          // :401-403
    void addBioSolver(const std::string& name);           // :405-408
    void addBioSolver(std::unique_ptr<BioSolver> b);
    // a BioSolver written against BioAbstractSolver::applyBioProcess(
    // GeometryHandler&, ForceSolver&); needs the GeometryHandler ctor
    void addBioSolver(std::unique_ptr<solver::BioAbstractSolver> b);
    void runSimulation();                                 // :69-110 (Parameters.iterations steps)
    void run(unsigned steps);
    void step();
    // the host table is brought up to date first, so that edits through it
    // extend what the device holds (they are uploaded before the next step)
    solver::ForceSolver& getForceSolver() {
        syncMirror();
        return forceSolver_;
    }
    geometry::Geometry& getGeometry() { return geo_; }
    lbib_ctx* context() { return ctx_; }
    GlobalSimulationParameters& parameters() { return params_; }
    int speciesIndex(const std::string& name) const;
    // host views: the device owns topology and forces between host edits;
    // syncMirror() copies them back, uploadTopology() pushes host edits.
    void syncMirror();
    void syncMarkersFromDevice() { syncMirror(); }
    void uploadTopology();
    void invalidateMirror() { mirrorValid_ = false; }
    // device-resident tension renewal (falls back to the host table when it
    // holds other force types)
    void renewTension(double force);
    // device-resident cell-junction renewal (tutorial_02_BioSolverCellJunction):
    // deleteForceType(7) and one spring pair per marker with a closest marker of
    // another cell within radius; k and l0 pass through the text descriptor
    void renewJunctions(double radius, double k, double l0);
    // the BioSolvers of the current iteration, outside a step
    void applyBioSolvers();
    // a registry BioSolver with its constants given: 3 Hookean tension (k, l0),
    // 4 growth (source), 5 threshold growth (maxsource, threshold),
    // 6 differentiation (threshold, species index), 7 / 8 cell division along
    // the longest axis / a random direction (max cell size), 9 PI area
    // regulator (target, target of cell 1, kp, maxsource)
    void addBioSolverWith(int kind, const double* prm, unsigned freq);
    // BioSolverHookeanMembraneTension: deleteForceType(6), then type-0 springs
    // (k, l0 through the text descriptor) along both directions of every connection
    void renewSprings(double k, double l0);
    // FluidSolver::addMass on every site of each domain d: table[d] (table[0]
    // for the extracellular sites), applied by the next collide
    void addMassByDomain(const std::vector<double>& table);
    // tutorial_02_BioSolverGrowth on the device (species tutorial_02_CDESolverD2Q5_R / _L)
    void thresholdGrowth(double threshold, double maxsource);
    // GeometryHandler::copyCellTypeToPhysicalNodes (GeometryHandler.cpp:267-301)
    // with the cell-type tracker map; domains missing from it are added as type 1
    void copyCellTypeToPhysicalNodes();
    // tutorial_01_BioSolverCellDivision / BioSolverCellDivision
    // (…CellDivision.cpp:42-256): every cell with an area above maxCellSize is
    // cut along the perpendicular bisector of its farthest node pair (or a
    // random direction through its centroid drawn from gen), the new ring
    // takes the next domain id, then cureLattice
    void divideCells(double maxCellSize, bool longestAxis, std::mt19937& gen);
    std::map<unsigned, double> computeAreas();            // GeometryHandler.cpp:101-134 (device shoelace)
    // GeometryHandler.cpp:136-147 (device per-domain sums, domains present on the lattice)
    std::map<unsigned, double> computeAccumulatedDomainConcentrations(const std::string& name);
    // GeometryHandler::cureLattice (GeometryHandler.cpp:149-170) after host
    // edits of the topology: upload, boundary nodes, scanline
    void cureLattice();
    // bumped whenever device state may have changed (a step enqueued or
    // resolved, an upload); host views re-read what they cache
    uint64_t stateGeneration() const { return gen_; }
    // waits for the enqueued iterations (a stop is resolved by run())
    void settleDevice() { settle(); }
    geometry::GeometryHandler* geometryHandler() { return gh_; }
    void detachGeometryHandler();
    const std::vector<std::string>& speciesNames() const { return speciesNames_; }
    StageTimes times;
    lbib_status lastStatus{};
    // Iterations enqueued before the device status is read (1 reads it after
    // every iteration, as the reference loop does).  A stop (perturbation,
    // pending remesh, capacity) is resolved from the iteration that raised it,
    // so results do not depend on this value.
    unsigned runAhead = 16;
    // Steps per recorded graph (0: every step is enqueued call by call).  A
    // multiple of 10 (the coarsening period) and of every BioSolver's period;
    // blocks start after iterations that are multiples of it.  A block that
    // cannot be recorded (a host-side BioSolver, an allocation or a status
    // read inside it) turns graph mode off and runs directly.
    void setGraphSteps(unsigned steps);
    unsigned graphSteps() const { return graphSteps_; }
    long graphLaunches = 0, graphCaptures = 0;
    std::string graphNote;   // why graph mode turned itself off
    // vtkFluidReporter / vtkCDEReporter / vtkCellReporter (ASCII VTK XML, no
    // VTK library): append the time step's block to <prefix>_<time>.vtm
    void writeVtkFluid(const std::string& prefix, unsigned time, int coarsening = 1);
    void writeVtkCDE(const std::string& prefix, unsigned time, int coarsening = 1);
    void writeVtkCells(const std::string& prefix, unsigned time);

 private:
    bool runGraphBlock(unsigned it0);
    void check(int rc, const char* what);
    void uploadForces();
    void handleGeometryStatus(unsigned it);
    void enqueueStep(unsigned it);
    void settle();
    void resolveStop(unsigned it);
    void hostTopologyFallback(unsigned it, bool grow);
    bool perturbFlagged(unsigned it);
    bool perturbRingsFlagged(unsigned it);
    std::string describe(const lbib_status& s);

    GlobalSimulationParameters params_;
    geometry::Geometry& geo_;
    int precision_, forceModel_;
    lbib_ctx* ctx_ = nullptr;
    std::vector<std::string> speciesNames_;
    std::map<std::string, lbib_species_spec> specs_;
    solver::ForceSolver forceSolver_;
    std::vector<std::unique_ptr<BioSolver>> bio_;
    std::vector<std::string> massSolvers_;
    std::vector<unsigned> ids_;                 // dense index -> node id (ascending)
    std::vector<std::vector<int>> bcDescriptors_;
    std::vector<std::map<std::string, std::vector<std::string>>> bcDescriptorText_;
    bool mirrorValid_ = true;
    // force types the device table may hold (bit t: type t); device-side
    // renewals that rewrite whole rows are valid only over their own type
    uint32_t devForceTypes_ = 0;
    size_t nm_ = 0, nc_ = 0;
    unsigned pending_ = 0;   // iterations enqueued since the last clear status
    // tutorial_02_CDESolverD2Q5_R/_L::initSolver draw (1 + U(-0.01, 0.01)) / 5
    // per distribution inside cells (…_R.cpp:44-58) from an unseeded device;
    // here the stream is seeded so that runs repeat
    std::mt19937 genInit_{1};
    std::vector<lbib_rect> rects_;
    unsigned graphSteps_ = 0;
    uint64_t gen_ = 0;
    geometry::GeometryHandler* gh_ = nullptr;
    const reportHandler::ReportHandler* report_ = nullptr;
};

namespace vtkxml {
// ASCII VTK XML without the VTK library (lbib_host_vtk.cpp)
void load_geometry(const std::string& file, geometry::Geometry& geo);
void write_image(const std::string& prefix, unsigned time, int nx, int ny, int coarse,
                 const std::vector<std::pair<std::string, std::pair<int, std::vector<double>>>>& fields);
void write_cells(const std::string& prefix, unsigned time, const geometry::Geometry& geo,
                 const std::map<unsigned, unsigned>& cellTypes);
void write_cells(const std::string& prefix, unsigned time,
                 const std::vector<std::shared_ptr<geometry::Connection>>& connections,
                 const std::map<unsigned, unsigned>& cellTypes);
}  // namespace vtkxml

}  // namespace LbmLib
