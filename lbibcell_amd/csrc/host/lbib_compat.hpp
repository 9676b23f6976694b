// lbib_compat.hpp -- the reference's plugin signatures over the device path.
//
// A BioSolver, reporter or main() written against LBIBCell's own interfaces
// compiles against these declarations:
//   * LbmLib::Parameters, the global GlobalSimulationParameters
//     (GlobalSimulationParameters.hpp:35-190);
//   * geometry::GeometryHandler(const Geometry&) with getPhysicalNodes(),
//     computeAreas(), computeAccumulatedDomainConcentrations(), cureLattice(),
//     copyCellTypeToPhysicalNodes() and getCellTypeTrackerMap()
//     (GeometryHandler.hpp:46-251);
//   * nodes::PhysicalNode with getDomainIdentifier(), getCellType(),
//     getFluidSolver() and getCDESolverSlow(name) (PhysicalNode.hpp:48-210);
//   * solver::BioAbstractSolver::applyBioProcess(GeometryHandler&, ForceSolver&)
//     and BioBaseSolver<T>, which registers T under T::name
//     (BioAbstractSolver.hpp:44-73, BioBaseSolver.hpp:45-96);
//   * reportHandler::ReportHandler and AbstractReportFunctor with the VTK
//     fluid, CDE and cell reporters (ReportHandler.hpp:34-65,
//     vtkFluidReporter.hpp:63-69, vtkCDEReporter.hpp:59-66,
//     vtkCellReporter.hpp:60-66);
//   * SimulationRunner(const GeometryHandler&, const ReportHandler&)
//     (SimulationRunner.hpp:51).
//
// The lattice lives on the GPU.  getPhysicalNodes() is a host view of it: a
// node reads domain ids, velocities, densities and concentrations from arrays
// copied from the device once per state change (the first read after a step
// copies the array it needs), and addMass is collected and applied to the
// device (lbib_add_mass_sites) when the BioSolver returns.  Citations are
// relative to the reference repository root; "LbmLib/" means libs/LbmLib/.
#pragma once
#include <atomic>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "lbib_host.hpp"

namespace LbmLib {

// UtilLib/include/Field.hpp: a pair of components
template <typename T>
struct Field {
    T x = T(), y = T();
    Field() = default;
    Field(T a, T b) : x(a), y(b) {}
};

// the reference's global parameter singleton (GlobalSimulationParameters.hpp:190)
extern GlobalSimulationParameters Parameters;

namespace geometry {
class GeometryHandler;
}
namespace nodes {
class PhysicalNode;
}

namespace detail {
class LatticeView;
}

namespace solver {

// FluidSolver of one site (FluidSolver.hpp:45-168), served from the view.
class FluidSolver {
 public:
    // FluidSolver::addMass (FluidSolver.cpp:253-257): overwrites the site's
    // mass for the next collide
    void addMass(double m);
    Field<double> getVelocity() const;   // FluidSolver::getVelocity (:242-246)
    double getRho() const;               // FluidSolver::getRho (:248-251)
    double getTau() const;

 private:
    friend class detail::LatticeView;
    detail::LatticeView* v_ = nullptr;
    long i_ = 0;
};

// CDESolver of one site and species (CDEAbstractSolver.hpp:39-176).
class CDEAbstractSolver {
 public:
    double getC() const;   // CDESolverD2Q5::getC (CDESolverD2Q5.cpp:71-73)
    std::string getName() const;
    double getTau() const;

 private:
    friend class detail::LatticeView;
    detail::LatticeView* v_ = nullptr;
    long i_ = 0;
    int species_ = 0;
};

// BioAbstractSolver (BioAbstractSolver.hpp:44-73)
class BioAbstractSolver {
 public:
    virtual ~BioAbstractSolver() {}
    virtual void applyBioProcess(geometry::GeometryHandler& geometryhandler, solver::ForceSolver& forcesolver) = 0;
    virtual std::string getName() = 0;

 protected:
    BioAbstractSolver() {}
};

// The name -> BioSolver factory (BioSolverFactory.hpp, Factory.hpp:105-160).
// Names are read when a solver is created, after static initialisation, so a
// T::name defined in another translation unit is ready by then.
class BioSolverFactory {
 public:
    using Creator = std::function<std::unique_ptr<BioAbstractSolver>()>;
    static bool add(std::function<std::string()> name, Creator create);
    // null when no registered solver has this name
    static std::unique_ptr<BioAbstractSolver> create(const std::string& name);
};

// BioBaseSolver<T> (BioBaseSolver.hpp:45-96): registers T under T::name.  As
// with the reference, T defines its default constructor (out of line), which
// instantiates the registration.
template <class T>
class BioBaseSolver : public BioAbstractSolver {
 public:
    std::string getName() override { return T::name; }

 protected:
    BioBaseSolver() { (void)registered_; }

 private:
    static bool registered_;
};
template <class T>
bool BioBaseSolver<T>::registered_ =
    BioSolverFactory::add([] { return std::string(T::name); },
                          [] { return std::unique_ptr<BioAbstractSolver>(new T()); });

}  // namespace solver

namespace nodes {

// PhysicalNode (PhysicalNode.hpp:48-210): one lattice site of the view.
class PhysicalNode {
 public:
    unsigned getDomainIdentifier() const;
    unsigned getCellType() const;
    int getXPos() const;
    int getYPos() const;
    solver::FluidSolver& getFluidSolver() { return fluid_; }
    detail::LatticeView* view() const { return v_; }
    // PhysicalNode::getCDESolverSlow (PhysicalNode.cpp:126-140): throws for
    // a name that is not registered
    solver::CDEAbstractSolver& getCDESolverSlow(const std::string& name) const;

 private:
    friend class detail::LatticeView;
    detail::LatticeView* v_ = nullptr;
    long i_ = 0;
    solver::FluidSolver fluid_;
};

}  // namespace nodes

namespace detail {

// Host copies of the device lattice for one GeometryHandler.  Each array is
// copied at its first read after the runner's state changed; addMass calls
// are collected until the BioSolver that made them returns.
class LatticeView {
 public:
    explicit LatticeView(geometry::GeometryHandler* gh) : gh_(gh), c_(LBIB_MAX_SPECIES) {
        for (auto& g : genC_) g = ~0ull;
    }
    const std::vector<std::vector<nodes::PhysicalNode*>>& grid();
    unsigned domain(long i);
    unsigned cellType(long i);
    Field<double> velocity(long i);
    double rho(long i);
    double conc(int species, long i);
    int species(const std::string& name) const;
    const std::string& speciesName(int k) const;
    double speciesTau(int k) const;
    double fluidTau() const;
    int nx() const { return nx_; }
    void addMass(long i, double m);
    // apply the collected addMass calls to the device, or drop them
    void flushMass();
    void discardMass();
    solver::CDEAbstractSolver& cde(int species, long i);
    SimulationRunner& runner() const;

 private:
    void shape();
    template <typename F>
    void refresh(std::atomic<uint64_t>& gen, F&& copy);
    geometry::GeometryHandler* gh_;
    int nx_ = 0, ny_ = 0;
    std::vector<nodes::PhysicalNode> nodes_;
    std::vector<std::vector<nodes::PhysicalNode*>> grid_;
    std::vector<std::vector<solver::CDEAbstractSolver>> cde_;
    std::vector<uint32_t> dom_;
    std::vector<double> ux_, uy_, rho_;
    std::vector<std::vector<double>> c_;
    std::atomic<uint64_t> genDom_{~0ull}, genVel_{~0ull}, genRho_{~0ull};
    std::atomic<uint64_t> genC_[LBIB_MAX_SPECIES];
    std::vector<int64_t> massSite_;
    std::vector<double> massVal_;
    std::mutex mu_;
};

}  // namespace detail

namespace geometry {

// GeometryHandler (GeometryHandler.hpp:46-251).  Constructed on a Geometry; the
// SimulationRunner built on it creates the device lattice and binds to it.
class GeometryHandler {
 public:
    explicit GeometryHandler(const Geometry& geometry);
    ~GeometryHandler();
    const std::vector<std::vector<nodes::PhysicalNode*>>& getPhysicalNodes() const;
    const Geometry& getGeometry() const { return geo_; }
    // GeometryHandler::computeAreas (GeometryHandler.cpp:101-134)
    const std::map<unsigned int, double> computeAreas() const;
    // GeometryHandler::computeAccumulatedDomainConcentrations (:136-147)
    std::map<unsigned int, double> computeAccumulatedDomainConcentrations(const std::string& name) const;
    // GeometryHandler::cureLattice (:149-170)
    void cureLattice();
    // GeometryHandler::copyCellTypeToPhysicalNodes (:267-313)
    void copyCellTypeToPhysicalNodes(std::map<unsigned int, unsigned int>& celltrackermap);
    void copyCellTypeToPhysicalNodes(unsigned int domainidentifier, unsigned int celltype);
    std::map<unsigned int, unsigned int>& getCellTypeTrackerMap(void);
    SimulationRunner& runner() const;
    // called by the runner around each BioSolver written against this class:
    // addMass calls left over from an interrupted call are dropped, those of
    // the finished call applied
    void beginBioSolver();
    void endBioSolver();

 private:
    friend class ::LbmLib::SimulationRunner;
    Geometry& geo_;
    SimulationRunner* runner_ = nullptr;
    std::unique_ptr<detail::LatticeView> view_;
};

}  // namespace geometry

namespace reportHandler {

// AbstractReportFunctor (AbstractReportFunctor.hpp:32-60)
class AbstractReportFunctor {
 public:
    explicit AbstractReportFunctor(const std::string& filename) : filename_(filename) {}
    virtual ~AbstractReportFunctor() {}
    virtual void operator()(unsigned int time) const = 0;

 protected:
    std::string filename_;
};

// ReportHandler (ReportHandler.hpp:34-65, ReportHandler.cpp)
class ReportHandler {
 public:
    explicit ReportHandler(unsigned int reportStep = 1) : reportStep_(reportStep ? reportStep : 1) {}
    void registerReporter(std::unique_ptr<AbstractReportFunctor> reporter) {
        reporters_.push_back(std::move(reporter));
    }
    void writeReport(unsigned int time) const {
        if (time % reportStep_ == 0)
            for (const auto& r : reporters_) (*r)(time);
    }
    void writeCrashReport(unsigned int time) const {
        for (const auto& r : reporters_) (*r)(time);
    }
    bool hasReporters() const { return !reporters_.empty(); }

 private:
    unsigned int reportStep_;
    std::vector<std::unique_ptr<AbstractReportFunctor>> reporters_;
};

// vtkFluidReporter (vtkFluidReporter.cpp:44-108): density and velocity of
// every coarsening-th site
class vtkFluidReporter : public AbstractReportFunctor {
 public:
    vtkFluidReporter(const std::vector<std::vector<nodes::PhysicalNode*>>& physicalNodes,
                     const unsigned int cdecoarseningfactor, const std::string& filename);
    void operator()(unsigned int time) const override;

 private:
    const std::vector<std::vector<nodes::PhysicalNode*>>& physicalNodes_;
    unsigned int coarse_;
};

// vtkCDEReporter (vtkCDEReporter.cpp:44-110): every species' concentration
class vtkCDEReporter : public AbstractReportFunctor {
 public:
    vtkCDEReporter(const std::vector<std::vector<nodes::PhysicalNode*>>& physicalNodes,
                   const unsigned int cdecoarseningfactor, const std::string& filename);
    void operator()(unsigned int time) const override;

 private:
    const std::vector<std::vector<nodes::PhysicalNode*>>& physicalNodes_;
    unsigned int coarse_;
};

// vtkCellReporter (vtkCellReporter.cpp:44-125): one polygon per cell with its
// cell type
class vtkCellReporter : public AbstractReportFunctor {
 public:
    vtkCellReporter(const std::vector<std::shared_ptr<geometry::Connection>>& connections,
                    const std::map<unsigned int, unsigned int>& celltypetrackermap, const std::string& filename);
    void operator()(unsigned int time) const override;

 private:
    const std::vector<std::shared_ptr<geometry::Connection>>& connections_;
    const std::map<unsigned int, unsigned int>& cellTypeTrackerMap_;
};

}  // namespace reportHandler

}  // namespace LbmLib
