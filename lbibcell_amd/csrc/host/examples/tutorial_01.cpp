// tutorial_01.cpp -- the main() sequence of the reference's apps/tutorial_01.cpp
// (apps/tutorial_01.cpp:38-152) on the device path, through the reference's
// own signatures (lbib_compat.hpp): Parameters from the parameter file, a
// Geometry and GeometryHandler, a ReportHandler with the VTK reporters, the
// SimulationRunner built on both, the force file, the box outlet and the five
// tutorial BioSolvers, then runSimulation().
//
//   tutorial_01 <parameters.txt> <geometry .vtm|.txt> <force.txt> <outdir>
//               [iterations] [registry|handler] [report]
//
// iterations overrides the file's Iterations; "handler" replaces the built-in
// tutorial_01_BioSolverAreaRegulator by AreaRegulatorOnHandler below, a
// BioSolver written against applyBioProcess(GeometryHandler&, ForceSolver&)
// that reads and writes the lattice through getPhysicalNodes(); "report"
// registers the VTK cell, CDE and fluid reporters at the file's ReportStep.
// After the run the state is written as raw arrays for comparison:
// <outdir>/f.bin (9 planes, natural direction order), g0.bin (5 planes),
// dom.bin (uint32), ids.bin (uint32), xy.bin (x, y pairs), conn.bin (first
// id, second id, domain as uint32 triples), plus the reference's text dump
// (writeSimulation) in <outdir>/{force,geo,parameter,solver}Out.txt.
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../lbib_compat.hpp"

namespace LbmLib {
namespace solver {

// Proportional area regulation (the scheme of
// tutorial_01_BioSolverAreaRegulator.cpp:43-100) written against the
// GeometryHandler: every cell aims at 200 lattice areas; its source is
// KP * (target - area) clamped to +-MAXSOURCE, and cells of type 1 grow at
// MAXSOURCE.
class AreaRegulatorOnHandler : public BioBaseSolver<AreaRegulatorOnHandler> {
 public:
    static const std::string name;
    AreaRegulatorOnHandler();

    void applyBioProcess(geometry::GeometryHandler& geometryhandler, ForceSolver& /*forcesolver*/) override {
        const unsigned every = 1;
        const double target = 200.0, kp = 0.00001, maxsource = 0.004;
        if (Parameters.getCurrentIteration() % every != 0) return;
        const std::map<unsigned, double> areas = geometryhandler.computeAreas();
        std::map<unsigned, double> source;
        for (const auto& a : areas) {
            if (targets_.count(a.first) == 0) targets_[a.first] = target;
            double s = kp * (targets_[a.first] - a.second);
            if (s > maxsource) s = maxsource;
            if (s < -maxsource) s = -maxsource;
            source[a.first] = s;
        }
        const auto& grid = geometryhandler.getPhysicalNodes();
        for (size_t y = 0; y < grid.size(); y++) {
            for (size_t x = 0; x < grid[y].size(); x++) {
                nodes::PhysicalNode* site = grid[y][x];
                const unsigned d = site->getDomainIdentifier();
                if (d == 0) continue;
                site->getFluidSolver().addMass(site->getCellType() == 1 ? maxsource : source[d]);
            }
        }
    }

 private:
    std::map<unsigned, double> targets_;
};
// defined out of line, as the reference's solvers are: this instantiates the
// BioBaseSolver<T> registration
AreaRegulatorOnHandler::AreaRegulatorOnHandler() : BioBaseSolver() {}
const std::string AreaRegulatorOnHandler::name = "AreaRegulatorOnHandler";

}  // namespace solver
}  // namespace LbmLib

namespace {
template <typename T>
void dump(const std::string& file, const std::vector<T>& v) {
    std::ofstream o(file, std::ios::binary);
    o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!o.good()) throw std::runtime_error("cannot write " + file);
}
void check(lbib_ctx* c, int rc) {
    if (rc != 0) throw std::runtime_error(lbib_last_error(c));
}
}  // namespace

int main(int argc, char* argv[]) {
    using namespace LbmLib;
    if (argc < 5) {
        std::cerr << "usage: tutorial_01 <parameters.txt> <geometry> <force.txt> <outdir> [iterations] "
                     "[registry|handler] [report]\n";
        return 2;
    }
    const std::string outfolder = std::string(argv[4]) + "/";
    const bool handler = argc > 6 && std::string(argv[6]) == "handler";
    const bool report = argc > 7 && std::string(argv[7]) == "report";
    try {
        mkdir(outfolder.c_str(), 0777);
        Parameters.loadGlobalSimulationParameters(argv[1]);
        if (argc > 5) Parameters.iterations = (unsigned)std::strtoul(argv[5], nullptr, 10);

        geometry::Geometry geo(argv[2]);
        geometry::GeometryHandler geohandler(geo);
        reportHandler::ReportHandler reporter(Parameters.getReportSteps());
        SimulationRunner simRunner(geohandler, reporter);

        if (report) {
            reporter.registerReporter(std::unique_ptr<reportHandler::AbstractReportFunctor>(
                new reportHandler::vtkCellReporter(geo.getConnections(), geohandler.getCellTypeTrackerMap(),
                                                   outfolder + "Cells")));
            reporter.registerReporter(std::unique_ptr<reportHandler::AbstractReportFunctor>(
                new reportHandler::vtkCDEReporter(geohandler.getPhysicalNodes(), 1, outfolder + "Cells")));
            reporter.registerReporter(std::unique_ptr<reportHandler::AbstractReportFunctor>(
                new reportHandler::vtkFluidReporter(geohandler.getPhysicalNodes(), 4, outfolder + "Cells")));
        }

        simRunner.initSolvers();
        simRunner.initForceSolver(argv[3]);

        simRunner.addMassSolver("MassSolverBoxOutlet");
        simRunner.addBioSolver(handler ? "AreaRegulatorOnHandler" : "tutorial_01_BioSolverAreaRegulator");
        simRunner.addBioSolver("tutorial_01_BioSolverMembraneTension");
        simRunner.addBioSolver("tutorial_01_BioSolverCellJunction");
        simRunner.addBioSolver("tutorial_01_BioSolverDifferentiation");
        simRunner.addBioSolver("tutorial_01_BioSolverCellDivision");

        simRunner.runSimulation();

        // the final state, for comparison
        simRunner.writeSimulation(outfolder + "forceOut.txt", outfolder + "geoOut.txt", outfolder + "parameterOut.txt",
                                  outfolder + "solverOut.txt");
        lbib_ctx* c = simRunner.context();
        const size_t n = (size_t)Parameters.getSizeX() * Parameters.getSizeY();
        std::vector<double> f(9 * n), g(5 * n);
        std::vector<uint32_t> dom(n);
        check(c, lbib_get_populations(c, f.data()));
        check(c, lbib_get_domains(c, dom.data()));
        dump(outfolder + "f.bin", f);
        dump(outfolder + "dom.bin", dom);
        if (!Parameters.getCdeSolvers().empty()) {
            check(c, lbib_get_species(c, 0, g.data()));
            dump(outfolder + "g0.bin", g);
        }
        simRunner.syncMirror();
        std::vector<uint32_t> ids, conn;
        std::vector<double> xy;
        for (const auto& kv : geo.getGeometryNodes()) {
            ids.push_back(kv.first);
            xy.push_back(kv.second->x);
            xy.push_back(kv.second->y);
        }
        for (const auto& cn : geo.getConnections()) {
            conn.push_back(cn->first->id);
            conn.push_back(cn->second->id);
            conn.push_back(cn->domain);
        }
        dump(outfolder + "ids.bin", ids);
        dump(outfolder + "xy.bin", xy);
        dump(outfolder + "conn.bin", conn);
        std::cout << "iterations " << Parameters.getCurrentIteration() << " markers " << ids.size() << std::endl;
    } catch (const std::exception& exp) {
        std::cout << exp.what() << std::endl;
        return 1;
    }
    return 0;
}
