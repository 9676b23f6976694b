// lbib_host.cpp -- host runtime (see lbib_host.hpp).
#include "lbib_host.hpp"
#include "lbib_compat.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <unordered_map>
#include <unordered_set>

namespace LbmLib {

namespace {
constexpr double kEPSILON = 10E-10;   // Constants.hpp:34
constexpr double kMAXLENGTH = 0.5;    // Constants.hpp:49
double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

// ------------------------------------------------------------- parameters --
void GlobalSimulationParameters::loadGlobalSimulationParameters(const std::string& file) {
    std::ifstream in(file);
    if (!in.is_open()) LBIB_FAIL("Cannot find the parameter file.");
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "CurrentIterations:") ls >> currentIteration;
        else if (key == "Iterations:") ls >> iterations;
        else if (key == "ReportStep:") ls >> reportSteps;
        else if (key == "SizeX:") ls >> sizeX;
        else if (key == "SizeY:") ls >> sizeY;
        else if (key == "tauFluid:") ls >> tauFluid;
        else if (key == "CDESolvers:") {
            std::string name;
            double tau;
            while (ls >> name >> tau) cdeSolvers[name] = tau;
        } else {
            throw Exception("An unknown parameter was passed in the global Parameter file");
        }
    }
}
// GlobalSimulationParameters::printParameters: the same fields as the file
void GlobalSimulationParameters::printParameters() const {
    std::cout << "CurrentIterations:\t" << currentIteration << "\nIterations:\t" << iterations << "\nReportStep:\t"
              << reportSteps << "\nSizeX:\t" << sizeX << "\nSizeY:\t" << sizeY << "\ntauFluid:\t" << tauFluid
              << "\nCDESolvers:";
    for (auto& kv : cdeSolvers) std::cout << '\t' << kv.first << '\t' << kv.second;
    std::cout << std::endl;
}
void GlobalSimulationParameters::writeGlobalSimulationParameters(const std::string& file) const {
    std::ofstream out(file);
    if (!out.is_open()) LBIB_FAIL("Cannot open the parameter output file.");
    out << "CurrentIterations:" << '\t' << currentIteration << "\n";
    out << "Iterations:" << '\t' << iterations << "\n";
    out << "ReportStep:" << '\t' << reportSteps << "\n";
    out << "SizeX:" << '\t' << sizeX << "\n";
    out << "SizeY:" << '\t' << sizeY << "\n";
    out << "tauFluid:" << '\t' << tauFluid << "\n";
    out << "CDESolvers:";
    for (auto& kv : cdeSolvers) out << '\t' << kv.first << '\t' << kv.second;
    out << "\n";
}

namespace geometry {

// ------------------------------------------------------------- geometry ----
Geometry::Geometry(const std::string& file) {
    auto ends = [&](const char* e) { return file.size() >= 4 && file.compare(file.size() - 4, 4, e) == 0; };
    if (ends(".txt")) loadGeometryTXT(file);
    else if (ends(".vtm") || ends(".vtp")) loadGeometryVTK(file);
    else LBIB_FAIL("geometry files are .txt, .vtm or .vtp");
}

void Geometry::loadGeometryVTK(const std::string& file) { vtkxml::load_geometry(file, *this); }

void Geometry::setBatchRects(const std::vector<lbib_rect>& rects) {
    rects_ = rects;
    rectStreams_.assign(rects.size(), Streams());
}

Geometry::Streams& Geometry::streamsAt(double x, double y) {
    for (size_t r = 0; r < rects_.size(); r++) {
        const lbib_rect& q = rects_[r];
        if (x >= q.x0 && x < q.x0 + q.w && y >= q.y0 && y < q.y0 + q.h) return rectStreams_[r];
    }
    return global_;
}

void Geometry::perturbatePosition(GeometryNode& n) {
    std::mt19937& gen = streamsAt(n.x, n.y).node;
    if (std::abs(static_cast<int>(n.x) - n.x) <= kEPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * kEPSILON) r = disNode_(gen);
        n.x = n.x + r;
        nodePerturbations++;
    }
    if (std::abs(static_cast<int>(n.y) - n.y) <= kEPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * kEPSILON) r = disNode_(gen);
        n.y = n.y + r;
        nodePerturbations++;
    }
    // The reference passes two draws as setPos(x + draw, y + draw); GCC
    // evaluates those call arguments right to left, so y takes the first draw.
    for (int pass = 0; pass < 2; pass++) {
        const double lhs = (pass == 0) ? (n.y - std::floor(n.y)) : (std::ceil(n.y) - n.y);
        if (std::abs(lhs - (n.x - std::floor(n.x))) <= kEPSILON) {
            const double ry = disNode_(gen);
            const double rx = disNode_(gen);
            n.x = n.x + rx;
            n.y = n.y + ry;
            nodePerturbations++;
        }
    }
}

void Geometry::setPos(GeometryNode& n, double x, double y) {
    n.x = x;
    n.y = y;
    perturbatePosition(n);
}

void Geometry::perturbConnection(Connection& c) {
    std::mt19937& genc = streamsAt(c.first->x, c.first->y).conn;
    if (std::abs(c.first->x - c.second->x) < kEPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * kEPSILON) r = disConn_(genc);
        setPos(*c.first, c.first->x + r, c.first->y);
        connectionPerturbations++;
    }
    if (std::abs(c.first->y - c.second->y) < kEPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * kEPSILON) r = disConn_(genc);
        setPos(*c.first, c.first->x, c.first->y + r);
        connectionPerturbations++;
    }
}

void Geometry::perturbConnections() {
    std::map<unsigned, Connection*> start;
    for (auto& c : connections_)
        if (!start.count(c->domain)) start[c->domain] = c.get();
    for (auto& kv : start) {
        Connection* c = kv.second;
        do {
            perturbConnection(*c);
            c = c->second->out;
            if (!c) LBIB_FAIL("open membrane ring");
        } while (c != kv.second);
    }
}

unsigned Geometry::addGeometryNode(unsigned id, double x, double y) {
    if (nodes_.count(id)) LBIB_FAIL("GeometryNode id already exists");
    auto n = std::unique_ptr<GeometryNode>(new GeometryNode());
    n->id = id;
    n->x = x;
    n->y = y;
    n->ox = x - 5.0;   // GeometryNode.cpp:102 forces the first stencil link
    n->oy = y - 5.0;
    perturbatePosition(*n);
    nodes_[id] = std::move(n);
    return id;
}
unsigned Geometry::addGeometryNode(double x, double y) {
    const unsigned id = nodes_.empty() ? 0u : nodes_.rbegin()->first + 1;
    return addGeometryNode(id, x, y);
}

void Geometry::addConnection(unsigned id1, unsigned id2, unsigned domain,
                             const std::map<std::string, std::vector<std::string>>& descriptor) {
    auto a = nodes_.find(id1), b = nodes_.find(id2);
    if (a == nodes_.end() || b == nodes_.end()) LBIB_FAIL("connection refers to an unknown GeometryNode");
    auto c = std::make_shared<Connection>();
    c->first = a->second.get();
    c->second = b->second.get();
    c->domain = domain;
    c->descriptor = descriptor;
    connections_.push_back(c);
    c->first->out = c.get();
    c->second->in = c.get();
    if (!cellTypes_.count(domain)) cellTypes_[domain] = 1;   // Geometry.cpp:140
}

void Geometry::eraseConnection(Connection* c) {
    if (c->first->out == c) c->first->out = nullptr;
    if (c->second->in == c) c->second->in = nullptr;
    connections_.erase(std::remove_if(connections_.begin(), connections_.end(),
                                      [c](const std::shared_ptr<Connection>& p) { return p.get() == c; }),
                       connections_.end());
}

void Geometry::removeGeometryNode(unsigned id) {
    auto it = nodes_.find(id);
    if (it == nodes_.end() || !it->second->in || !it->second->out)
        LBIB_FAIL("Geometry::removeGeometryNode: Could not remove the GeometryNode");
    Connection* c1 = it->second->in;
    Connection* c2 = it->second->out;
    GeometryNode* n1 = c1->first;
    GeometryNode* n2 = c2->second;
    const auto descriptor = c1->descriptor;
    const unsigned dom = c1->domain;
    eraseConnection(c1);
    eraseConnection(c2);
    nodes_.erase(it);
    addConnection(n1->id, n2->id, dom, descriptor);
}

unsigned Geometry::remeshBoundary(double maxLength) {
    // Each long connection is erased (order of the others kept) and its two
    // halves appended in visiting order; a survivor list does it in one pass.
    std::vector<std::shared_ptr<Connection>> copy(connections_), keep, added;
    keep.reserve(copy.size());
    unsigned counter = 0;
    for (auto& c : copy) {
        const double dx = c->first->x - c->second->x, dy = c->first->y - c->second->y;
        if (std::sqrt(dx * dx + dy * dy) > maxLength) {
            const double x1 = c->first->x, y1 = c->first->y, x2 = c->second->x, y2 = c->second->y;
            const unsigned nid = addGeometryNode((x1 + x2) / 2, (y1 + y2) / 2);
            GeometryNode* m = nodes_.at(nid).get();
            GeometryNode* a = c->first;
            GeometryNode* b = c->second;
            if (a->out == c.get()) a->out = nullptr;
            if (b->in == c.get()) b->in = nullptr;
            for (int h = 0; h < 2; h++) {
                auto n = std::make_shared<Connection>();
                n->first = h == 0 ? a : m;
                n->second = h == 0 ? m : b;
                n->domain = c->domain;
                n->descriptor = c->descriptor;
                n->first->out = n.get();
                n->second->in = n.get();
                added.push_back(n);
            }
            counter++;
        } else {
            keep.push_back(c);
        }
    }
    if (counter) {
        keep.insert(keep.end(), added.begin(), added.end());
        connections_.swap(keep);
    }
    return counter;
}

unsigned Geometry::coarsenBoundary(double maxLength) {
    std::vector<std::shared_ptr<Connection>> all(connections_);
    std::unordered_map<const Connection*, bool> dead;
    unsigned removed = 0;
    for (auto it = nodes_.begin(); it != nodes_.end();) {
        GeometryNode& n = *it->second;
        const GeometryNode* p1 = n.in->first;
        const GeometryNode* p2 = n.out->second;
        const double len = std::sqrt((p2->x - p1->x) * (p2->x - p1->x) + (p2->y - p1->y) * (p2->y - p1->y));
        if (len < maxLength) {
            Connection* c1 = n.in;
            Connection* c2 = n.out;
            GeometryNode* n1 = c1->first;
            GeometryNode* n2 = c2->second;
            auto nc = std::make_shared<Connection>();
            nc->first = n1;
            nc->second = n2;
            nc->domain = c1->domain;
            nc->descriptor = c1->descriptor;
            for (Connection* c : {c1, c2}) {
                if (c->first->out == c) c->first->out = nullptr;
                if (c->second->in == c) c->second->in = nullptr;
                dead[c] = true;
            }
            it = nodes_.erase(it);
            n1->out = nc.get();
            n2->in = nc.get();
            all.push_back(nc);
            removed++;
        } else {
            ++it;
        }
    }
    if (removed) {
        std::vector<std::shared_ptr<Connection>> out;
        out.reserve(all.size());
        for (auto& c : all)
            if (!dead.count(c.get())) out.push_back(c);
        connections_.swap(out);
    }
    return removed;
}

void Geometry::rebuild(const std::vector<unsigned>& ids, const std::vector<double>& xy, const std::vector<double>& vxy,
                       const std::vector<double>& oxy, const std::vector<int32_t>& a, const std::vector<int32_t>& b,
                       const std::vector<uint32_t>& dom,
                       const std::vector<std::map<std::string, std::vector<std::string>>>& descriptors) {
    nodes_.clear();
    connections_.clear();
    std::vector<GeometryNode*> byIndex(ids.size());
    for (size_t k = 0; k < ids.size(); k++) {
        auto n = std::unique_ptr<GeometryNode>(new GeometryNode());
        n->id = ids[k];
        n->x = xy[2 * k]; n->y = xy[2 * k + 1];
        n->vx = vxy[2 * k]; n->vy = vxy[2 * k + 1];
        n->ox = oxy[2 * k]; n->oy = oxy[2 * k + 1];
        byIndex[k] = n.get();
        nodes_[ids[k]] = std::move(n);
    }
    connections_.reserve(a.size());
    for (size_t j = 0; j < a.size(); j++) {
        auto c = std::make_shared<Connection>();
        c->first = byIndex[a[j]];
        c->second = byIndex[b[j]];
        c->domain = dom[j];
        if (j < descriptors.size()) c->descriptor = descriptors[j];
        c->first->out = c.get();
        c->second->in = c.get();
        connections_.push_back(c);
    }
}

bool Geometry::checkGeometryIntegrity() const {
    for (auto& kv : nodes_) {
        const GeometryNode* n = kv.second.get();
        if (!n->in || !n->out || n->in->second != n || n->out->first != n) return false;
    }
    return connections_.size() == nodes_.size();
}

void Geometry::loadGeometryTXT(const std::string& file) {
    std::ifstream in(file);
    if (!in.is_open()) LBIB_FAIL("Cannot find the geometry.txt file.");
    if (!nodes_.empty() || !connections_.empty()) LBIB_FAIL("Only load the geometry once.");
    std::string line;
    bool readingNodes = true;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line.find("#Connection") != std::string::npos) readingNodes = false;
        if (line[0] == '#') continue;
        std::stringstream ls(line);
        if (readingNodes) {
            unsigned id;
            double x, y;
            ls >> id >> x >> y;
            addGeometryNode(id, x, y);
        } else {
            unsigned id1, id2, dom;
            ls >> id1 >> id2 >> dom;
            std::map<std::string, std::vector<std::string>> desc;
            std::string b, s;
            while (ls >> b >> s) desc[b].push_back(s);
            addConnection(id1, id2, dom, desc);
            cellTypes_[dom] = 1;   // all cell types are 1 for txt geometries
        }
    }
}

void Geometry::writeGeometry(const std::string& file) const {
    std::ofstream out(file);
    if (!out.is_open()) LBIB_FAIL("Cannot open the output file for the geometry.");
    out << "#Nodes (id\txPos\tyPos)\n";
    for (auto& kv : nodes_) out << kv.first << '\t' << kv.second->x << '\t' << kv.second->y << '\n';
    out << "#Connection (nodeId1\tnodeId2\tdomainId\tbsolver\tcdesolver\t...)\n";
    for (auto& c : connections_) {
        out << c->first->id << '\t' << c->second->id << '\t' << c->domain;
        for (auto& kv : c->descriptor)
            for (auto& s : kv.second) out << '\t' << kv.first << '\t' << s;
        out << '\n';
    }
}

}  // namespace geometry

namespace solver {

void ForceSolver::addForce(std::stringstream* d) {
    std::stringstream copy(d->str());
    unsigned type, id;
    copy >> type >> id;
    ForceStruct f{};
    f.type = (int)type;
    f.id1 = id;
    std::stringstream s(d->str());
    unsigned t;
    s >> t;
    switch (type) {
        case 0: case 7: s >> f.id1 >> f.id2 >> f.p[0] >> f.p[1]; break;
        case 5: case 6: s >> f.id1 >> f.id2 >> f.p[0]; break;
        case 1: s >> f.id1 >> f.p[0] >> f.p[1] >> f.p[2] >> f.p[3]; break;
        case 2: s >> f.id1 >> f.p[0] >> f.p[1]; break;
        case 3: case 4: s >> f.id1 >> f.p[0] >> f.p[1] >> f.p[2]; break;
        default: LBIB_FAIL("Invalid force solver type found.");
    }
    addForce(f);
}

void ForceSolver::loadForceFile(const std::string& file) {
    std::ifstream in(file);
    if (!in.is_open()) LBIB_FAIL("Cannot find the force solver file.");
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ls(line);
        addForce(&ls);
    }
}

void ForceSolver::writeForceSolver(const std::string& file) const {
    std::ofstream out(file);
    if (!out.is_open()) throw Exception("Problem to open the output file for the force solver");
    for (auto& kv : forcemap_)
        for (auto& f : kv.second) {
            out << f.type << '\t' << f.id1;
            switch (f.type) {
                case 0: case 7: out << '\t' << f.id2 << '\t' << f.p[0] << '\t' << f.p[1]; break;
                case 5: case 6: out << '\t' << f.id2 << '\t' << f.p[0]; break;
                case 1: out << '\t' << f.p[0] << '\t' << f.p[1] << '\t' << f.p[2] << '\t' << f.p[3]; break;
                case 2: out << '\t' << f.p[0] << '\t' << f.p[1]; break;
                default: out << '\t' << f.p[0] << '\t' << f.p[1] << '\t' << f.p[2]; break;
            }
            out << '\n';
        }
}

bool ForceSolver::onlyType(int t) const {
    for (auto& kv : forcemap_)
        for (auto& f : kv.second)
            if (f.type != t) return false;
    return true;
}

void ForceSolver::deleteForceType(unsigned type) {
    for (auto& kv : forcemap_)
        kv.second.erase(std::remove_if(kv.second.begin(), kv.second.end(),
                                       [type](const ForceStruct& f) { return (unsigned)f.type == type; }),
                        kv.second.end());
    dirty_ = true;
}

}  // namespace solver

// ------------------------------------------------------------ bio solvers --
namespace {

// tutorial_01_BioSolverMembraneTension (tutorial_01_BioSolverMembraneTension.cpp:42-77;
// tutorial_02_: FORCE 0.005, BioSolverMembraneTension: FORCE 0.02):
// every FREQUENCY iterations, type-6 tension of FORCE along both directions of
// every connection.  FORCE goes through a text descriptor (6 significant digits).
struct MembraneTension : BioSolver {
    unsigned freq = 10;
    double force = 0.01;
    explicit MembraneTension(double f = 0.01) : force(f) {}
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        r.renewTension(force);
    }
};

// tutorial_02_BioSolverCellJunction, tutorial_01_BioSolverCellJunction and
// BioSolverCellJunction (tutorial_02_BioSolverCellJunction.cpp:37-99): every
// FREQUENCY = 10 iterations, junction springs K = 0.02, L0 = 0.5 between each
// node and its closest node of another cell within RADIUS = 1.0.
struct CellJunction : BioSolver {
    unsigned freq = 10;
    double radius = 1.0, k = 0.02, l0 = 0.5;
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        r.renewJunctions(radius, k, l0);
    }
};

// BioSolverHookeanMembraneTension (BioSolverHookeanMembraneTension.cpp:34-80):
// every 10 iterations type-0 springs K = 0.01, l_resting = 0.05.
struct HookeanTension : BioSolver {
    unsigned freq = 10;
    double k = 0.01, l0 = 0.05;
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        r.renewSprings(k, l0);
    }
};

// BioSolverGrowth (BioSolverGrowth.cpp:33-66): addMass(0.0025) on every cell site.
struct Growth : BioSolver {
    unsigned freq = 1;
    double source = 0.0025;
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        unsigned maxdom = 0;
        for (auto& c : r.getGeometry().getConnections()) maxdom = std::max(maxdom, c->domain);
        for (auto& kv : r.getGeometry().cellTypeTrackerMap()) maxdom = std::max(maxdom, kv.first);
        std::vector<double> t(maxdom + 1, source);
        t[0] = 0.0;
        r.addMassByDomain(t);
    }
};

// tutorial_02_BioSolverGrowth (tutorial_02_BioSolverGrowth.cpp:33-82): one device op.
struct ThresholdGrowth : BioSolver {
    unsigned freq = 1;
    double maxsource = 0.001, threshold = 1.0;
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        r.thresholdGrowth(threshold, maxsource);
    }
};

// BioSolverAreaRegulator (BioSolverAreaRegulator.cpp:33-130): PI form with
// KI = 0, cell 1 aiming at TARGETAREA2 = 1000 before iteration 5000, sources
// clamped to [0, MAXSOURCE]; cell type 1 sites take MAXSOURCE.
struct AreaRegulatorPI : BioSolver {
    unsigned freq = 1;
    double target = 200.0, target2 = 1000.0, kp = 0.00001, maxsource = 0.004;
    std::map<unsigned, double> targets, integ;
    void applyBioProcess(SimulationRunner& r) override {
        const unsigned it = r.parameters().currentIteration;
        if (it % freq != 0) return;
        const double KI = 0.0;
        if (it < 5000) { targets[1] = target2; integ[1] = 0.0; }
        else targets[1] = target;
        auto areas = r.computeAreas();
        for (auto& kv : areas)
            if (targets.find(kv.first) == targets.end()) { targets[kv.first] = target; integ[kv.first] = 0.0; }
        std::map<unsigned, double> err, src;
        for (auto& kv : areas) err[kv.first] = targets[kv.first] - areas[kv.first];
        for (auto& kv : areas) integ[kv.first] += err[kv.first];
        for (auto& kv : areas) {
            const double u = kp * err[kv.first] + KI * integ[kv.first];
            src[kv.first] = (u > maxsource) ? maxsource : ((u < 0.0) ? 0.0 : u);
        }
        auto& ctm = r.getGeometry().cellTypeTrackerMap();
        unsigned maxdom = 0;
        for (auto& kv : areas) maxdom = std::max(maxdom, kv.first);
        for (auto& kv : ctm) maxdom = std::max(maxdom, kv.first);
        std::vector<double> t(maxdom + 1, 0.0);
        for (unsigned d = 1; d <= maxdom; d++) {
            auto c = ctm.find(d);
            const unsigned type = (c == ctm.end()) ? 1u : c->second;   // copyCellTypeToPhysicalNodes default
            auto sv = src.find(d);
            t[d] = (type == 1) ? maxsource : (sv == src.end() ? 0.0 : sv->second);
        }
        r.addMassByDomain(t);
    }
};

// tutorial_01_BioSolverDifferentiation / BioSolverDifferentiation
// (…Differentiation.cpp:33-83): every 100 iterations, cells whose accumulated
// concentration (GeometryHandler.cpp:136-147) is below THRESHOLD = 1.6 become
// cell type 2.  The tracker is reset once per solver (the reference's
// std::call_once flag is shared by all instances of its class).
struct Differentiation : BioSolver {
    unsigned freq = 100;
    double threshold = 1.6;
    std::string species = "tutorial_01_CDESolverD2Q5_SIGNAL";
    bool inited = false;
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        auto acc = r.computeAccumulatedDomainConcentrations(species);
        auto& ctm = r.getGeometry().cellTypeTrackerMap();
        if (!inited) {
            std::map<unsigned, int> cells;
            for (auto& c : r.getGeometry().getConnections()) cells[c->domain] = 1;
            ctm.clear();
            ctm[0] = 0;
            for (auto& kv : cells) ctm[kv.first] = 1;
            inited = true;
        }
        if (!acc.empty())
            for (auto it = ++acc.begin(); it != acc.end(); ++it)
                if (it->second < threshold) ctm[it->first] = 2;
        r.copyCellTypeToPhysicalNodes();
    }
};

// tutorial_01_/tutorial_02_BioSolverCellDivision (longest axis) and
// BioSolverCellDivision (random direction): every 100 iterations, cells above
// MAXCELLSIZE = 380 divide.  The random direction draws from this solver's
// mt19937(1) (the reference's generator is a function-static shared by all
// instances of the class).
struct CellDivision : BioSolver {
    unsigned freq = 100;
    double maxsize = 380.0;
    bool longest = true;
    std::mt19937 gen{1};
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        r.divideCells(maxsize, longest, gen);
    }
};

// BioSolverEmpty (BioSolverEmpty.cpp:37-48)
struct Empty : BioSolver {
    void applyBioProcess(SimulationRunner&) override { std::cout << "I'm BioSolverEmpty" << std::endl; }
    unsigned graphPeriod() const override { return 0; }
};

// tutorial_01_BioSolverAreaRegulator (tutorial_01_BioSolverAreaRegulator.cpp:43-100)
// served by one device op (areas, clamped source, addMass on every cell site).
struct AreaRegulator : BioSolver {
    unsigned freq = 1;
    double target = 200.0, kp = 0.00001, maxsource = 0.004, factor = 0.0;
    unsigned graphPeriod() const override { return freq; }
    void applyBioProcess(SimulationRunner& r) override {
        if (r.parameters().currentIteration % freq != 0) return;
        if (lbib_area_regulator(r.context(), target, kp, maxsource, factor) != 0)
            throw Exception(lbib_last_error(r.context()));
    }
};

}  // namespace

// --------------------------------------------------------- the runner ------
SimulationRunner::SimulationRunner(const GlobalSimulationParameters& params, geometry::Geometry& geo, int precision,
                                   int forceModel)
    : params_(params), geo_(geo), precision_(precision), forceModel_(forceModel) {
    for (auto& kv : params_.cdeSolvers) {
        if (!knownCDESolver(kv.first)) throw std::bad_typeid();   // Factory.hpp:45-47
        speciesNames_.push_back(kv.first);
        lbib_species_spec s;
        std::memset(&s, 0, sizeof(s));
        s.tau = kv.second;
        if (kv.first == "CDESolverD2Q5") {
            s.kind = LBIB_SPECIES_PLAIN;
        } else if (kv.first == "tutorial_02_CDESolverD2Q5_R" || kv.first == "tutorial_02_CDESolverD2Q5_L") {
            // …_R.cpp:38-42 / …_L.cpp:38-42: gamma, a = 0.1 or b = 0.9, deltaT;
            // the partner index is resolved below once every name is known
            const bool isR = kv.first.back() == 'R';
            s.kind = isR ? LBIB_SPECIES_SCHNAKENBERG_R : LBIB_SPECIES_SCHNAKENBERG_L;
            s.production = 100.0;
            s.source_inside = isR ? 0.1 : 0.9;
            s.decay = 1.0e-4;
            s.init_inside = 1.0;
            s.prod_domain = -1;
        } else {
            // tutorial_01_CDESolverD2Q5_SIGNAL constants (…_SIGNAL.cpp:39-43)
            s.kind = LBIB_SPECIES_SIGNAL;
            s.production = 0.0001;
            s.decay = 0.00001;
            s.switchoff = 5000;
            s.init_inside = 1.0;
            s.prod_domain = 1;
        }
        specs_[kv.first] = s;
    }
    // getCDESolverSlow(partner) (…_R.cpp:97, …_L.cpp:96) throws when the
    // partner is not registered; the lookup is done once here
    for (auto& kv : specs_) {
        if (kv.second.kind != LBIB_SPECIES_SCHNAKENBERG_R && kv.second.kind != LBIB_SPECIES_SCHNAKENBERG_L) continue;
        const std::string partner = kv.second.kind == LBIB_SPECIES_SCHNAKENBERG_R ? "tutorial_02_CDESolverD2Q5_L"
                                                                                  : "tutorial_02_CDESolverD2Q5_R";
        const int j = speciesIndex(partner);
        if (j < 0) LBIB_FAIL("The CDE Solver with the name " + partner + " does not exists.");
        kv.second.prod_domain = j;
    }
}

SimulationRunner::~SimulationRunner() {
    if (gh_ && gh_->runner_ == this) gh_->runner_ = nullptr;
    if (ctx_) lbib_destroy(ctx_);
}

bool SimulationRunner::knownCDESolver(const std::string& name) {
    return name == "CDESolverD2Q5" || name == "tutorial_01_CDESolverD2Q5_SIGNAL" ||
           name == "tutorial_02_CDESolverD2Q5_R" || name == "tutorial_02_CDESolverD2Q5_L" ||
           name.rfind("SIGNAL", 0) == 0 || name.rfind("LINEAR", 0) == 0;
}

void SimulationRunner::setSpeciesSpec(const std::string& name, const lbib_species_spec& spec) {
    if (!specs_.count(name)) LBIB_FAIL("The CDE Solver with the name " + name + " does not exists.");
    const double tau = specs_[name].tau;
    specs_[name] = spec;
    specs_[name].tau = tau;
}

int SimulationRunner::speciesIndex(const std::string& name) const {
    for (size_t k = 0; k < speciesNames_.size(); k++)
        if (speciesNames_[k] == name) return (int)k;
    return -1;
}

void SimulationRunner::check(int rc, const char* what) {
    if (rc != 0) throw Exception(std::string(what) + ": " + lbib_last_error(ctx_));
}

void SimulationRunner::setup() {
    lbib_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.nx = (int)params_.sizeX;
    cfg.ny = (int)params_.sizeY;
    cfg.tau_fluid = params_.tauFluid;
    cfg.force_model = forceModel_;
    cfg.precision = precision_;
    cfg.nspecies = (int)speciesNames_.size();
    if (cfg.nspecies > LBIB_MAX_SPECIES) LBIB_FAIL("at most 4 CDE species per context");
    for (int k = 0; k < cfg.nspecies; k++) cfg.species[k] = specs_[speciesNames_[k]];
    cfg.max_markers = (int)geo_.getGeometryNodes().size() * 8 + 1024;
    cfg.max_connections = cfg.max_markers;
    if (lbib_create(&cfg, &ctx_) != 0) {
        std::string m = ctx_ ? lbib_last_error(ctx_) : "lbib_create failed";
        if (ctx_) lbib_destroy(ctx_);
        ctx_ = nullptr;
        throw Exception(m);
    }
    if (!rects_.empty()) check(lbib_set_batch_layout(ctx_, (int)rects_.size(), rects_.data()), "batch layout");
    // GeometryHandler::cureLattice: perturb connections, link stencils (implicit
    // on the device), boundary nodes, scanline.
    geo_.perturbConnections();
    uploadTopology();
    check(lbib_cure_lattice(ctx_), "cureLattice");
    check(lbib_read_status(ctx_, &lastStatus), "status");
    if (lastStatus.bits & (LBIB_ERR_MASK | LBIB_HALT_MASK)) throw Exception("cureLattice: " + describe(lastStatus));
    check(lbib_clear_status(ctx_), "status");
}

void SimulationRunner::initSolvers(const double* u0) {
    check(lbib_init_solvers(ctx_, u0), "initSolvers");
    gen_++;
    std::vector<uint32_t> dom;
    std::vector<double> g;
    const size_t n = (size_t)params_.sizeX * params_.sizeY;
    std::uniform_real_distribution<> dis(-0.01, 0.01);
    for (size_t k = 0; k < speciesNames_.size(); k++) {
        const lbib_species_spec& s = specs_[speciesNames_[k]];
        if (s.kind != LBIB_SPECIES_SCHNAKENBERG_R && s.kind != LBIB_SPECIES_SCHNAKENBERG_L) continue;
        if (dom.empty()) {
            dom.resize(n);
            check(lbib_get_domains(ctx_, dom.data()), "initSolvers");
        }
        // …_R.cpp:44-58: 0 outside cells, init_inside * (1 + U(-0.01, 0.01)) / 5
        // per distribution inside, sites in (y, x) order
        g.resize(5 * n);
        check(lbib_get_species(ctx_, (int)k, g.data()), "initSolvers");
        for (size_t i = 0; i < n; i++) {
            if (dom[i] == 0) continue;
            for (int q = 0; q < 5; q++) g[q * n + i] = s.init_inside * (1.0 + dis(genInit_)) / 5.0;
        }
        check(lbib_set_species(ctx_, (int)k, g.data()), "initSolvers");
    }
}

namespace {
// The stored state of the reference is post-collision with localSwap applied
// (FluidSolver.cpp:259-264, CDESolverD2Q5.cpp:151-154): slot q of a site holds
// the natural-order value of the opposite direction (T stays).
constexpr int kSlotOfD2Q9[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr int kSlotOfD2Q5[5] = {0, 3, 4, 1, 2};

// std::ostream << double with the default precision and floatfield: %g
inline void put_g(std::string& out, double v) {
    char b[32];
    const int k = std::snprintf(b, sizeof(b), "%g", v);
    out.append(b, (size_t)k);
}
}  // namespace

void SimulationRunner::writeSimulation(const std::string& forceFile, const std::string& geometryFile,
                                       const std::string& parameterFile, const std::string& solverFile) {
    syncMirror();
    forceSolver_.writeForceSolver(forceFile);
    geo_.writeGeometry(geometryFile);
    params_.writeGlobalSimulationParameters(parameterFile);
    std::ofstream out(solverFile);
    if (!out.is_open()) LBIB_FAIL("Cannot open the output file for the force solver.");
    const int nx = (int)params_.sizeX, ny = (int)params_.sizeY;
    const size_t n = (size_t)nx * ny;
    auto dump = [&](const std::string& name, const std::vector<double>& nat, int nq, const int* slot_of) {
        out << "#BEGIN " << name << "\n";
        std::string line;
        for (int y = 0; y < ny; y++) {
            for (int x = 0; x < nx; x++) {
                line.clear();
                line += std::to_string(x);
                line += '\t';
                line += std::to_string(y);
                const size_t i = (size_t)y * nx + x;
                for (int q = 0; q < nq; q++) {
                    line += '\t';
                    put_g(line, nat[(size_t)slot_of[q] * n + i]);
                }
                line += '\n';
                out << line;
            }
        }
        out << "#END " << name << "\n";
    };
    out << "#Solvers (xPos\tyPos\tDistributions...)\n";
    std::vector<double> f(9 * n);
    check(lbib_get_populations(ctx_, f.data()), "get_populations");
    dump("FluidSolver", f, 9, kSlotOfD2Q9);
    std::vector<double> g(5 * n);
    for (size_t k = 0; k < speciesNames_.size(); k++) {
        check(lbib_get_species(ctx_, (int)k, g.data()), "get_species");
        dump(speciesNames_[k], g, 5, kSlotOfD2Q5);
    }
    if (!out.good()) LBIB_FAIL("Cannot open the output file for the force solver.");
}

void SimulationRunner::initSolvers(const std::string& solverFile) {
    gen_++;
    std::ifstream in(solverFile);
    if (!in.is_open()) LBIB_FAIL("Cannot find the force file.");
    const int nx = (int)params_.sizeX, ny = (int)params_.sizeY;
    const size_t n = (size_t)nx * ny;
    // start from the current state: sites the file omits keep their values
    std::vector<double> f(9 * n);
    check(lbib_get_populations(ctx_, f.data()), "get_populations");
    std::vector<std::vector<double>> g(speciesNames_.size(), std::vector<double>(5 * n));
    for (size_t k = 0; k < g.size(); k++) check(lbib_get_species(ctx_, (int)k, g[k].data()), "get_species");
    std::string line, solverName;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::stringstream ls(line);
        if (line[0] != '#') {
            int x, y;
            ls >> x >> y;
            if (x < 0 || y < 0 || x >= nx || y >= ny) LBIB_FAIL("The position does not match");
            const size_t i = (size_t)y * nx + x;
            if (solverName == "FluidSolver") {
                for (int q = 0; q < 9; q++) ls >> f[(size_t)kSlotOfD2Q9[q] * n + i];
            } else {
                const int k = speciesIndex(solverName);
                if (k < 0) LBIB_FAIL("The CDE Solver with the name " + solverName + " does not exists.");
                for (int q = 0; q < 5; q++) ls >> g[k][(size_t)kSlotOfD2Q5[q] * n + i];
            }
        } else if (line.find("#BEGIN") != std::string::npos) {
            std::string temp;
            ls >> temp >> solverName;
        } else if (line.find("#END") != std::string::npos) {
            solverName.clear();
        }
    }
    check(lbib_set_populations(ctx_, f.data()), "set_populations");
    for (size_t k = 0; k < g.size(); k++) check(lbib_set_species(ctx_, (int)k, g[k].data()), "set_species");
}

void SimulationRunner::initForceSolver(const std::string& file) { forceSolver_.loadForceFile(file); }

void SimulationRunner::addMassSolver(const std::string& name) {
    if (name != "MassSolverBoxOutlet") throw std::bad_typeid();
    massSolvers_.push_back(name);
    check(lbib_set_box_outlet(ctx_, 1), "addMassSolver");
}

void SimulationRunner::addBioSolver(const std::string& name) {
    // a BioSolver registered through BioBaseSolver<T> (BioSolverFactory)
    if (gh_) {
        auto user = solver::BioSolverFactory::create(name);
        if (user) {
            addBioSolver(std::move(user));
            return;
        }
    }
    std::unique_ptr<BioSolver> b;
    if (name == "tutorial_01_BioSolverMembraneTension") b.reset(new MembraneTension(0.01));
    else if (name == "tutorial_02_BioSolverMembraneTension") b.reset(new MembraneTension(0.005));
    else if (name == "BioSolverMembraneTension") b.reset(new MembraneTension(0.02));
    else if (name == "BioSolverHookeanMembraneTension") b.reset(new HookeanTension());
    else if (name == "tutorial_01_BioSolverAreaRegulator") b.reset(new AreaRegulator());
    else if (name == "BioSolverAreaRegulator") b.reset(new AreaRegulatorPI());
    else if (name == "tutorial_02_BioSolverCellJunction" || name == "tutorial_01_BioSolverCellJunction" ||
             name == "BioSolverCellJunction")
        b.reset(new CellJunction());
    else if (name == "BioSolverGrowth") b.reset(new Growth());
    else if (name == "tutorial_02_BioSolverGrowth") b.reset(new ThresholdGrowth());
    else if (name == "tutorial_01_BioSolverDifferentiation") b.reset(new Differentiation());
    else if (name == "BioSolverDifferentiation") {
        auto* d = new Differentiation();
        d->species = "CDESolverD2Q5HH";
        b.reset(d);
    } else if (name == "tutorial_01_BioSolverCellDivision" || name == "tutorial_02_BioSolverCellDivision")
        b.reset(new CellDivision());
    else if (name == "BioSolverCellDivision") {
        auto* d = new CellDivision();
        d->longest = false;
        b.reset(d);
    } else if (name == "BioSolverEmpty") b.reset(new Empty());
    else throw std::bad_typeid();
    b->name = name;
    bio_.push_back(std::move(b));
}

void SimulationRunner::addBioSolver(std::unique_ptr<BioSolver> b) { bio_.push_back(std::move(b)); }

void SimulationRunner::addBioSolverWith(int kind, const double* prm, unsigned freq) {
    if (freq == 0) LBIB_FAIL("frequency must be positive");
    std::unique_ptr<BioSolver> b;
    switch (kind) {
        case 3: {
            auto* h = new HookeanTension();
            h->k = prm[0]; h->l0 = prm[1]; h->freq = freq;
            h->name = "BioSolverHookeanMembraneTension";
            b.reset(h);
            break;
        }
        case 4: {
            auto* g = new Growth();
            g->source = prm[0]; g->freq = freq;
            g->name = "BioSolverGrowth";
            b.reset(g);
            break;
        }
        case 5: {
            auto* g = new ThresholdGrowth();
            g->maxsource = prm[0]; g->threshold = prm[1]; g->freq = freq;
            g->name = "tutorial_02_BioSolverGrowth";
            b.reset(g);
            break;
        }
        case 6: {
            const int k = (int)prm[1];
            if (k < 0 || k >= (int)speciesNames_.size()) LBIB_FAIL("differentiation: species index out of range");
            auto* d = new Differentiation();
            d->threshold = prm[0]; d->species = speciesNames_[k]; d->freq = freq;
            d->name = "tutorial_01_BioSolverDifferentiation";
            b.reset(d);
            break;
        }
        case 7:
        case 8: {
            auto* d = new CellDivision();
            d->maxsize = prm[0]; d->longest = kind == 7; d->freq = freq;
            d->name = kind == 7 ? "tutorial_01_BioSolverCellDivision" : "BioSolverCellDivision";
            b.reset(d);
            break;
        }
        case 9: {
            auto* a = new AreaRegulatorPI();
            a->target = prm[0]; a->target2 = prm[1]; a->kp = prm[2]; a->maxsource = prm[3]; a->freq = freq;
            a->name = "BioSolverAreaRegulator";
            b.reset(a);
            break;
        }
        default: throw std::bad_typeid();
    }
    bio_.push_back(std::move(b));
}

void SimulationRunner::syncMirror() {
    if (mirrorValid_) return;
    settle();
    const size_t nm = nm_, nc = nc_;
    std::vector<unsigned> ids(nm);
    std::vector<double> xy(2 * nm), v(2 * nm), o(2 * nm);
    check(lbib_get_marker_ids(ctx_, ids.data()), "get_marker_ids");
    check(lbib_get_markers(ctx_, xy.data(), v.data(), nullptr, o.data()), "get_markers");
    std::vector<int32_t> a(nc), b(nc), bc(nc);
    std::vector<uint32_t> dom(nc);
    check(lbib_get_connections(ctx_, a.data(), b.data(), dom.data(), bc.data()), "get_connections");
    std::vector<std::map<std::string, std::vector<std::string>>> desc(nc);
    for (size_t j = 0; j < nc; j++)
        if (bc[j] >= 0) desc[j] = bcDescriptorText_.at(bc[j]);
    geo_.rebuild(ids, xy, v, o, a, b, dom, desc);
    ids_ = ids;
    // force table back into ForceSolver::forcemap_ (ids instead of indices)
    std::vector<int32_t> off(nm + 1);
    check(lbib_get_forces(ctx_, off.data(), nullptr, nullptr, nullptr), "get_forces");
    const int nf = off[nm];
    std::vector<int32_t> type(nf), partner(nf);
    std::vector<double> prm(4 * (size_t)nf);
    check(lbib_get_forces(ctx_, off.data(), type.data(), partner.data(), prm.data()), "get_forces");
    std::map<unsigned, std::vector<solver::ForceStruct>> fm;
    for (size_t k = 0; k < nm; k++)
        for (int j = off[k]; j < off[k + 1]; j++) {
            solver::ForceStruct f{type[j], ids[k], partner[j] >= 0 ? ids[partner[j]] : 0u,
                                  {prm[4 * j], prm[4 * j + 1], prm[4 * j + 2], prm[4 * j + 3]}};
            fm[ids[k]].push_back(f);
        }
    forceSolver_.replace(std::move(fm));
    forceSolver_.dirty_ = false;
    mirrorValid_ = true;
}

void SimulationRunner::renewTension(double force) {
    // FORCE travels through a text descriptor in the reference: 6 significant digits
    std::stringstream ss;
    ss << force;
    double f;
    ss >> f;
    if (forceSolver_.dirty_ || !forceSolver_.onlyType(6) || (devForceTypes_ & ~(1u << 6))) {
        syncMirror();
        forceSolver_.deleteForceType(6);
        for (auto& c : geo_.getConnections()) {
            forceSolver_.addForce(solver::ForceStruct{6, c->first->id, c->second->id, {f, 0, 0, 0}});
            forceSolver_.addForce(solver::ForceStruct{6, c->second->id, c->first->id, {f, 0, 0, 0}});
        }
        uploadForces();
        return;
    }
    check(lbib_forces_tension(ctx_, f), "tension");
    devForceTypes_ |= 1u << 6;
    mirrorValid_ = false;   // the host forcemap no longer matches the device table
}

void SimulationRunner::renewJunctions(double radius, double k, double l0) {
    // K and L0 travel through a text descriptor in the reference
    auto via_text = [](double v) {
        std::stringstream ss;
        ss << v;
        double r;
        ss >> r;
        return r;
    };
    if (forceSolver_.dirty_) {
        syncMirror();
        uploadForces();
    }
    check(lbib_forces_junctions(ctx_, radius, via_text(k), via_text(l0)), "cell junctions");
    devForceTypes_ |= 1u << 7;
    mirrorValid_ = false;
}

void SimulationRunner::uploadTopology() {
    settle();
    gen_++;
    const auto& nodes = geo_.getGeometryNodes();
    const int nm = (int)nodes.size();
    ids_.clear();
    std::unordered_map<const geometry::GeometryNode*, int> index;
    std::vector<double> xy(2 * nm), v(2 * nm), o(2 * nm);
    int k = 0;
    for (auto& kv : nodes) {
        const auto& n = *kv.second;
        ids_.push_back(kv.first);
        index[&n] = k;
        xy[2 * k] = n.x; xy[2 * k + 1] = n.y;
        v[2 * k] = n.vx; v[2 * k + 1] = n.vy;
        o[2 * k] = n.ox; o[2 * k + 1] = n.oy;
        k++;
    }
    check(lbib_set_markers(ctx_, nm, xy.data(), v.data(), o.data()), "set_markers");
    check(lbib_set_marker_ids(ctx_, ids_.data()), "set_marker_ids");
    // boundary-condition descriptors (BoundaryNode.cpp:43-46)
    std::map<std::vector<int>, int> descIndex;
    bcDescriptors_.clear();
    bcDescriptorText_.clear();
    const auto& conns = geo_.getConnections();
    const int nc = (int)conns.size();
    std::vector<int32_t> a(nc), b(nc), bc(nc, -1);
    std::vector<uint32_t> dom(nc);
    uint32_t maxdom = 0;
    for (int c = 0; c < nc; c++) {
        a[c] = index.at(conns[c]->first);
        b[c] = index.at(conns[c]->second);
        dom[c] = conns[c]->domain;
        maxdom = std::max(maxdom, dom[c]);
        for (auto& kv : conns[c]->descriptor) {
            if (kv.first != "BoundarySolverNoFluxD2Q5") throw std::bad_typeid();
            std::vector<int> sp;
            for (auto& s : kv.second) {
                const int si = speciesIndex(s);
                if (si < 0) throw Exception("The CDE Solver with the name " + s + " does not exists.");
                sp.push_back(si);
            }
            auto it = descIndex.find(sp);
            if (it == descIndex.end()) {
                it = descIndex.emplace(sp, (int)bcDescriptors_.size()).first;
                bcDescriptors_.push_back(sp);
                bcDescriptorText_.push_back(conns[c]->descriptor);
            }
            bc[c] = it->second;
        }
    }
    if (!bcDescriptors_.empty()) {
        std::vector<int32_t> off{0}, sp;
        for (auto& d : bcDescriptors_) {
            for (int s : d) sp.push_back(s);
            off.push_back((int32_t)sp.size());
        }
        check(lbib_set_bc_descriptors(ctx_, (int)bcDescriptors_.size(), off.data(), sp.data()), "set_bc");
    }
    check(lbib_set_connections(ctx_, nc, a.data(), b.data(), dom.data(), bc.data()), "set_connections");
    // cell types per domain id (missing domains read as 1, GeometryHandler.cpp:278-284)
    auto& ctm = geo_.cellTypeTrackerMap();
    for (auto& kv : ctm) maxdom = std::max(maxdom, kv.first);
    std::vector<uint32_t> ct(maxdom + 1, 1u);
    for (auto& kv : ctm) ct[kv.first] = kv.second;
    check(lbib_set_celltypes(ctx_, (int)ct.size(), ct.data()), "set_celltypes");
    nm_ = (size_t)nm;
    nc_ = (size_t)nc;
    uploadForces();
    mirrorValid_ = true;
}

void SimulationRunner::uploadForces() {
    settle();
    const int nm = (int)ids_.size();
    std::vector<int32_t> off(nm + 1, 0), type, partner;
    std::vector<double> prm;
    auto dense = [&](unsigned id) -> int {
        auto it = std::lower_bound(ids_.begin(), ids_.end(), id);
        if (it == ids_.end() || *it != id) throw Exception("Could not compute and set the forces.");
        return (int)(it - ids_.begin());
    };
    const auto& fm = forceSolver_.forcemap();
    for (auto& kv : fm)
        if (!kv.second.empty()) dense(kv.first);
    for (int k = 0; k < nm; k++) {
        auto it = fm.find(ids_[k]);
        if (it != fm.end()) {
            for (auto& f : it->second) {
                type.push_back(f.type);
                const bool pair = f.type == 0 || f.type == 5 || f.type == 6 || f.type == 7;
                partner.push_back(pair ? dense(f.id2) : -1);
                for (int j = 0; j < 4; j++) prm.push_back(f.p[j]);
            }
        }
        off[k + 1] = (int32_t)type.size();
    }
    check(lbib_set_forces(ctx_, nm, off.data(), type.data(), partner.data(), prm.data()), "set_forces");
    devForceTypes_ = 0;
    for (int t : type) devForceTypes_ |= 1u << t;
    forceSolver_.dirty_ = false;
}

std::string SimulationRunner::describe(const lbib_status& s) {
    std::string m;
    if (s.bits & LBIB_ERR_VELOCITY) m += "The magnitude of velocity is too high; ";
    if (s.bits & LBIB_ERR_EDGE) m += "geometry point moved too close to the boundary; ";
    if (s.bits & LBIB_ERR_SCANLINE) m += "domainID at the left boundary must be 0.; ";
    if (s.bits & LBIB_ERR_SLOPE) m += "This can only happen if something went wrong with the perturbation of the connection.; ";
    if (s.bits & LBIB_ERR_CAPACITY) m += "device work list capacity exceeded; ";
    if (s.bits & LBIB_ERR_TOPOLOGY) m += "a domain is not one closed membrane ring of at most 4096 markers; ";
    char buf[160];
    snprintf(buf, sizeof(buf), "[status bits 0x%x, markers %d, connections %d, iteration %u] ", s.bits, s.n_markers,
             s.n_connections, params_.currentIteration);
    m += buf;
    // the device's ids: remeshing there may have renumbered the dense indices
    // since the host last copied the topology
    if (s.first_bad_marker >= 0 && s.first_bad_marker < s.n_markers) {
        std::vector<uint32_t> ids((size_t)s.n_markers);
        if (lbib_get_marker_ids(ctx_, ids.data()) == 0)
            m += "GeometryNode ID=" + std::to_string(ids[(size_t)s.first_bad_marker]);
    }
    return m;
}

// Remeshing/coarsening on the host mirror, used when the device halts on a
// perturbation draw or runs out of capacity (both rare).
void SimulationRunner::hostTopologyFallback(unsigned it, bool grow) {
    auto t0 = std::chrono::steady_clock::now();
    check(lbib_clear_status(ctx_), "status");
    mirrorValid_ = false;
    syncMirror();
    while (geo_.remeshBoundary(kMAXLENGTH)) {}
    if (it % 10 == 0) {
        forceSolver_.deleteAllForces();
        geo_.coarsenBoundary(kMAXLENGTH);
    }
    if (grow) {
        const int nm = (int)geo_.getGeometryNodes().size(), nc = (int)geo_.getConnections().size();
        check(lbib_reserve(ctx_, 2 * nm + 1024, 2 * nc + 1024), "reserve");
    }
    uploadTopology();
    times.host += seconds_since(t0);
}

// Geometry::moveGeometryNodes -> setPos -> perturbatePosition draws only for
// the nodes that meet one of its conditions, in id order: fetch just those
// from the device, perturb them on the host and write their positions back,
// then redo stage B.  Returns false (nothing changed) when more nodes are
// flagged than one fetch holds; the caller then takes the whole-geometry path.
bool SimulationRunner::perturbFlagged(unsigned it) {
    constexpr int kCap = 4096;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int32_t> idx(kCap);
    std::vector<uint32_t> ids(kCap);
    std::vector<double> xy(2 * kCap);
    int cnt = 0;
    check(lbib_perturbation_candidates(ctx_, kCap, &cnt, idx.data(), ids.data(), xy.data()), "perturbation candidates");
    if (cnt > kCap) return false;
    std::vector<int> order(cnt);
    for (int i = 0; i < cnt; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return ids[a] < ids[b]; });
    std::vector<int32_t> at(cnt);
    std::vector<double> pos(2 * cnt);
    for (int j = 0; j < cnt; j++) {
        const int i = order[j];
        geometry::GeometryNode n;
        n.id = ids[i];
        n.x = xy[2 * i];
        n.y = xy[2 * i + 1];
        geo_.perturbatePosition(n);
        at[j] = idx[i];
        pos[2 * j] = n.x;
        pos[2 * j + 1] = n.y;
    }
    check(lbib_set_marker_positions(ctx_, cnt, at.data(), pos.data()), "set_marker_positions");
    mirrorValid_ = false;
    check(lbib_clear_status(ctx_), "status");
    times.sparsePerturbations++;
    times.host += seconds_since(t0);
    check(lbib_step_geometry(ctx_, it, it % 10 == 0), "moveLattice");
    return true;
}

// GeometryHandler::perturbConnections walks every ring from its first
// connection, in increasing domain id, and draws only for connections whose
// markers line up within kEPSILON (a draw moves the connection's first marker,
// which can line up the ring's last connection in turn).  Rings share no
// markers, so only the rings the device flags need the walk: fetch those in
// walk order, walk them on the host and write back the markers that moved.
// Returns false (nothing changed) when they do not fit one fetch.
bool SimulationRunner::perturbRingsFlagged(unsigned it) {
    constexpr int kRings = 1024, kNodes = 1 << 17;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> dom(kRings);
    std::vector<int32_t> off(kRings + 1), mk(kNodes);
    std::vector<double> xy(2 * (size_t)kNodes);
    int nr = 0, nn = 0;
    check(lbib_perturbation_rings(ctx_, kRings, kNodes, &nr, &nn, dom.data(), off.data(), mk.data(), xy.data()),
          "perturbation rings");
    if (nr > kRings || nn > kNodes) return false;
    std::vector<int32_t> at;
    std::vector<double> pos;
    for (int r = 0; r < nr; r++) {
        const int o = off[r], n = off[r + 1] - off[r];
        std::vector<geometry::GeometryNode> ring(n);
        for (int j = 0; j < n; j++) {
            ring[j].x = xy[2 * (size_t)(o + j)];
            ring[j].y = xy[2 * (size_t)(o + j) + 1];
        }
        for (int j = 0; j < n; j++) {
            geometry::Connection cn;
            cn.first = &ring[j];
            cn.second = &ring[(j + 1) % n];
            cn.domain = dom[r];
            geo_.perturbConnection(cn);
        }
        for (int j = 0; j < n; j++) {
            if (ring[j].x == xy[2 * (size_t)(o + j)] && ring[j].y == xy[2 * (size_t)(o + j) + 1]) continue;
            at.push_back(mk[o + j]);
            pos.push_back(ring[j].x);
            pos.push_back(ring[j].y);
        }
    }
    check(lbib_set_marker_positions(ctx_, (int)at.size(), at.data(), pos.data()), "set_marker_positions");
    mirrorValid_ = false;
    check(lbib_clear_status(ctx_), "status");
    times.sparsePerturbations++;
    times.host += seconds_since(t0);
    check(lbib_step_geometry(ctx_, it, it % 10 == 0), "moveLattice");
    return true;
}

void SimulationRunner::handleGeometryStatus(unsigned it) {
    bool coarsened = false;
    for (int guard = 0; guard < 64; guard++) {
        check(lbib_read_status(ctx_, &lastStatus), "status");
        nm_ = (size_t)lastStatus.n_markers;
        nc_ = (size_t)lastStatus.n_connections;
        const uint32_t bits = lastStatus.bits;
        if (bits & LBIB_ERR_MASK) throw Exception(describe(lastStatus));
        if ((bits & LBIB_HALT_PERTURB_NODE) && !(bits & LBIB_HALT_PERTURB_CONN) && perturbFlagged(it)) continue;
        if ((bits & LBIB_HALT_PERTURB_CONN) && !(bits & LBIB_HALT_PERTURB_NODE) && perturbRingsFlagged(it)) continue;
        if (bits & (LBIB_HALT_PERTURB_NODE | LBIB_HALT_PERTURB_CONN)) {
            // the move ran, nothing after it: perturb on the host, redo stage B
            auto t0 = std::chrono::steady_clock::now();
            times.fullPerturbations++;
            mirrorValid_ = false;
            syncMirror();
            if (bits & LBIB_HALT_PERTURB_NODE) {
                // Geometry::moveGeometryNodes -> setPos -> perturbatePosition, id order
                for (auto& kv : geo_.nodes()) geo_.perturbatePosition(*kv.second);
            } else {
                geo_.perturbConnections();
            }
            std::vector<double> xy(2 * nm_), v(2 * nm_), o(2 * nm_);
            size_t k = 0;
            for (auto& kv : geo_.getGeometryNodes()) {
                xy[2 * k] = kv.second->x; xy[2 * k + 1] = kv.second->y;
                v[2 * k] = kv.second->vx; v[2 * k + 1] = kv.second->vy;
                o[2 * k] = kv.second->ox; o[2 * k + 1] = kv.second->oy;
                k++;
            }
            check(lbib_set_markers(ctx_, (int)nm_, xy.data(), v.data(), o.data()), "set_markers");
            check(lbib_clear_status(ctx_), "status");
            times.host += seconds_since(t0);
            check(lbib_step_geometry(ctx_, it, it % 10 == 0), "moveLattice");
            continue;
        }
        if (bits & (LBIB_HALT_PERTURB_REMESH | LBIB_HALT_GROW)) {
            hostTopologyFallback(it, (bits & LBIB_HALT_GROW) != 0);
            coarsened = true;
            break;
        }
        if (bits & LBIB_FLAG_REMESH) {
            // more remesh rounds than one stage-B pass runs (then coarsen)
            check(lbib_clear_status(ctx_), "status");
            check(lbib_step_topology(ctx_, it % 10 == 0), "remesh");
            continue;
        }
        break;
    }
    (void)coarsened;
    if (it % 10 == 0) {
        // ForceSolver::deleteAllForces before the BioSolvers renew them
        forceSolver_.deleteAllForces();
        forceSolver_.dirty_ = false;
        check(lbib_forces_clear(ctx_), "deleteAllForces");
        devForceTypes_ = 0;
    }
    check(lbib_clear_status(ctx_), "status");
}

namespace {
// Raised by settle() when enqueued iterations hit a stop; run() resolves it.
struct RunAheadStop {
    unsigned it;
};
constexpr uint32_t kStopBits = LBIB_ERR_MASK | LBIB_HALT_MASK | LBIB_FLAG_REMESH | LBIB_SKIPPED;
}  // namespace

// One SimulationRunner iteration (SimulationRunner.cpp:72-100) enqueued
// without reading the status: lattice stage, moveLattice + remesh/coarsen,
// deleteAllForces every 10th iteration, BioSolvers.
void SimulationRunner::enqueueStep(unsigned it) {
    params_.currentIteration = it;
    gen_++;
    auto t0 = std::chrono::steady_clock::now();
    check(lbib_step_begin(ctx_, it), "step begin");
    pending_++;
    if (forceSolver_.dirty_) {
        syncMirror();
        uploadForces();
    }
    check(lbib_step_lattice(ctx_, it), "lattice step");
    auto t1 = std::chrono::steady_clock::now();
    check(lbib_step_geometry(ctx_, it, it % 10 == 0), "moveLattice");
    mirrorValid_ = false;
    if (it % 10 == 0) {
        // ForceSolver::deleteAllForces before the BioSolvers renew them
        forceSolver_.deleteAllForces();
        forceSolver_.dirty_ = false;
        check(lbib_forces_clear(ctx_), "deleteAllForces");
        devForceTypes_ = 0;
    }
    auto t2 = std::chrono::steady_clock::now();
    times.lattice += std::chrono::duration<double>(t1 - t0).count();
    times.geometry += std::chrono::duration<double>(t2 - t1).count();
    // mass solvers: MassSolverBoxOutlet is applied inside the collide kernel.
    for (auto& b : bio_) b->applyBioProcess(*this);
    times.bio += seconds_since(t2);
    params_.currentIteration = it + 1;
    times.steps++;
}

// Reads the status of everything enqueued; throws RunAheadStop when an
// enqueued iteration stopped.  Host reads and uploads of the topology call it
// first, so they never act on (or overwrite) state past a stop.
void SimulationRunner::settle() {
    if (pending_ == 0) return;
    auto t0 = std::chrono::steady_clock::now();
    check(lbib_read_status(ctx_, &lastStatus), "status");
    times.wait += seconds_since(t0);
    nm_ = (size_t)lastStatus.n_markers;
    nc_ = (size_t)lastStatus.n_connections;
    if (lastStatus.bits & kStopBits) throw RunAheadStop{lastStatus.step_iteration};
    pending_ = 0;
}

// Iteration it stopped and nothing after it changed device state: restore the
// launch state of iteration it + 1, finish it the way the synchronous loop
// does, and drop host-side edits made while enqueueing the skipped iterations.
void SimulationRunner::resolveStop(unsigned it) {
    check(lbib_rewind(ctx_, it + 1), "rewind");
    gen_++;
    pending_ = 0;
    times.stops++;
    params_.currentIteration = it;
    mirrorValid_ = false;
    forceSolver_.dirty_ = false;
    handleGeometryStatus(it);
    auto t2 = std::chrono::steady_clock::now();
    for (auto& b : bio_) b->applyBioProcess(*this);
    times.bio += seconds_since(t2);
    params_.currentIteration = it + 1;
}

void SimulationRunner::step() { run(1); }

void SimulationRunner::renewSprings(double k, double l0) {
    auto via_text = [](double v) {
        std::stringstream ss;
        ss << v;
        double r;
        ss >> r;
        return r;
    };
    const double K = via_text(k), L = via_text(l0);
    if (forceSolver_.dirty_ || devForceTypes_ != 0) {
        syncMirror();
        forceSolver_.deleteForceType(6);
        for (auto& c : geo_.getConnections()) {
            forceSolver_.addForce(solver::ForceStruct{0, c->first->id, c->second->id, {K, L, 0, 0}});
            forceSolver_.addForce(solver::ForceStruct{0, c->second->id, c->first->id, {K, L, 0, 0}});
        }
        uploadForces();
        return;
    }
    check(lbib_forces_springs(ctx_, K, L), "Hookean tension");
    devForceTypes_ |= 1u;
    mirrorValid_ = false;
}

void SimulationRunner::addMassByDomain(const std::vector<double>& table) {
    check(lbib_add_mass_by_domain(ctx_, (int)table.size(), table.data()), "addMass");
}

void SimulationRunner::thresholdGrowth(double threshold, double maxsource) {
    const int r = speciesIndex("tutorial_02_CDESolverD2Q5_R"), l = speciesIndex("tutorial_02_CDESolverD2Q5_L");
    if (r < 0) LBIB_FAIL("The CDE Solver with the name tutorial_02_CDESolverD2Q5_R does not exists.");
    if (l < 0) LBIB_FAIL("The CDE Solver with the name tutorial_02_CDESolverD2Q5_L does not exists.");
    check(lbib_mass_growth(ctx_, r, l, threshold, maxsource), "growth");
}

void SimulationRunner::copyCellTypeToPhysicalNodes() {
    auto& ctm = geo_.cellTypeTrackerMap();
    unsigned maxdom = 0;
    for (auto& c : geo_.getConnections()) {
        maxdom = std::max(maxdom, c->domain);
        if (!ctm.count(c->domain)) ctm[c->domain] = 1;
    }
    for (auto& kv : ctm) maxdom = std::max(maxdom, kv.first);
    std::vector<uint32_t> ct(maxdom + 1, 1u);
    for (auto& kv : ctm) ct[kv.first] = kv.second;
    settle();
    check(lbib_set_celltypes(ctx_, (int)ct.size(), ct.data()), "copyCellTypeToPhysicalNodes");
}

namespace {
using CellDef = std::map<unsigned, std::vector<std::shared_ptr<geometry::Connection>>>;
CellDef cellDefinition(const geometry::Geometry& g) {   // updateCellDefinition (…CellDivision.cpp:250-256)
    CellDef d;
    for (auto& c : g.getConnections()) d[c->domain].push_back(c);
    return d;
}
bool cutByLine(const geometry::Connection& c, double invslope, double intercept) {   // (…CellDivision.cpp:184-190)
    const bool C1 = c.first->y > invslope * c.first->x + intercept;
    const bool C2 = c.second->y < invslope * c.second->x + intercept;
    return (C1 && C2) || (!C1 && !C2);
}
// getTwoConnectionsLongestAxis (…CellDivision.cpp:151-198)
std::vector<std::shared_ptr<geometry::Connection>> longestAxisCut(
    const std::vector<std::shared_ptr<geometry::Connection>>& cell) {
    double dist = 0.0;
    const geometry::GeometryNode *N1 = nullptr, *N2 = nullptr;
    for (auto& c1 : cell)
        for (auto& c2 : cell) {
            // boost::geometry::distance (pythagoras): a sum of two squares
            const double dx = c1->first->x - c2->first->x, dy = c1->first->y - c2->first->y;
            const double t = std::sqrt(dx * dx + dy * dy);
            if (t > dist) { dist = t; N1 = c1->first; N2 = c2->first; }
        }
    if (!N1) LBIB_FAIL("could not find cell division plane");
    const double mx = 0.5 * (N1->x + N2->x), my = 0.5 * (N1->y + N2->y);
    const double invslope = -(N2->x - N1->x) / (N2->y - N1->y);
    const double intercept = my - invslope * mx;
    if (!std::isfinite(invslope) || !std::isfinite(intercept)) LBIB_FAIL("could not find cell division plane");
    std::vector<std::shared_ptr<geometry::Connection>> out;
    for (auto& c : cell)
        if (cutByLine(*c, invslope, intercept)) out.push_back(c);
    if (out.size() != 2) LBIB_FAIL("More or less than 2 connections found in BioSolver::getTwoConnectionsLongestAxis().");
    return out;
}
// getTwoConnectionsRandomDirection (…CellDivision.cpp:200-248)
std::vector<std::shared_ptr<geometry::Connection>> randomDirectionCut(
    const std::vector<std::shared_ptr<geometry::Connection>>& cell, std::mt19937& gen) {
    std::uniform_real_distribution<> dis(0.0, 2.0 * M_PI);
    double cx = 0.0, cy = 0.0;
    for (auto& c : cell) { cx += c->first->x; cy += c->first->y; }
    cx /= (double)cell.size();
    cy /= (double)cell.size();
    std::vector<std::shared_ptr<geometry::Connection>> out;
    unsigned counter = 0;
    while (out.size() != 2) {
        const double invslope = std::tan(dis(gen));
        const double intercept = cy - invslope * cx;
        out.clear();
        for (auto& c : cell)
            if (cutByLine(*c, invslope, intercept)) out.push_back(c);
        counter++;
        if (counter > 10) LBIB_FAIL("could not find cell division plane");
    }
    return out;
}
}  // namespace

void SimulationRunner::divideCells(double maxCellSize, bool longestAxis, std::mt19937& gen) {
    syncMirror();
    auto& ctm = geo_.cellTypeTrackerMap();
    CellDef def = cellDefinition(geo_);
    if (ctm.empty()) {
        ctm[0] = 0;
        for (auto& kv : def) ctm[kv.first] = 1;
    }
    auto areas = computeAreas();
    areas[0] = 0;
    bool divided = false;
    for (auto it = ++areas.begin(); it != areas.end(); ++it) {
        if (!(it->second > maxCellSize)) continue;
        // divideCell (…CellDivision.cpp:94-149)
        const unsigned dom = it->first;
        const unsigned newID = def.rbegin()->first + 1;
        auto aff = longestAxis ? longestAxisCut(def[dom]) : randomDirectionCut(def[dom], gen);
        aff[0]->first->out = nullptr; aff[0]->second->in = nullptr;
        aff[1]->first->out = nullptr; aff[1]->second->in = nullptr;
        if (aff[0]->descriptor != aff[1]->descriptor)
            LBIB_FAIL("The method cannot handle cell division cases with heterogeneous boundary conditions.");
        const auto desc = aff[0]->descriptor;
        geo_.addConnection(aff[0]->first->id, aff[1]->second->id, dom, desc);
        geo_.addConnection(aff[1]->first->id, aff[0]->second->id, dom, desc);
        for (auto& c : aff) geo_.eraseConnection(c.get());
        geometry::Connection* start = aff[1]->second->out;
        geometry::Connection* c1 = start;
        geometry::Connection* c2 = nullptr;
        start->domain = newID;
        while (c2 != start) {
            c2 = c1->second->out;
            c2->domain = newID;
            c1 = c2;
        }
        def = cellDefinition(geo_);
        ctm[def.rbegin()->first] = ctm[dom];
        geo_.perturbConnections();   // GeometryHandler::cureLattice step 1
        divided = true;
    }
    if (divided) {
        // the rest of cureLattice (boundary nodes, scanline) on the device
        uploadTopology();
        check(lbib_cure_lattice(ctx_), "cureLattice");
        check(lbib_read_status(ctx_, &lastStatus), "status");
        if (lastStatus.bits & (LBIB_ERR_MASK | LBIB_HALT_MASK)) throw Exception("cureLattice: " + describe(lastStatus));
        check(lbib_clear_status(ctx_), "status");
    }
    copyCellTypeToPhysicalNodes();
}

void SimulationRunner::cureLattice() {
    syncMirror();
    geo_.perturbConnections();
    uploadTopology();
    check(lbib_cure_lattice(ctx_), "cureLattice");
    check(lbib_read_status(ctx_, &lastStatus), "status");
    if (lastStatus.bits & (LBIB_ERR_MASK | LBIB_HALT_MASK)) throw Exception("cureLattice: " + describe(lastStatus));
    check(lbib_clear_status(ctx_), "status");
    gen_++;
}

void SimulationRunner::applyBioSolvers() {
    settle();
    for (auto& b : bio_) b->applyBioProcess(*this);
}

void SimulationRunner::setGraphSteps(unsigned steps) {
    if (steps % 10 != 0) LBIB_FAIL("graph blocks must be a multiple of the coarsening period (10)");
    graphSteps_ = steps;
    if (ctx_) check(lbib_graph_reset(ctx_), "graph reset");
}

// One block of graphSteps_ iterations from it0 (it0 - 1 a multiple of the
// block length): replayed from the recorded graph, recorded first when there
// is none or the launch state changed.  false: nothing of the block ran (graph
// mode is off from now on) and the caller enqueues it step by step.
bool SimulationRunner::runGraphBlock(unsigned it0) {
    const unsigned B = graphSteps_;
    gen_++;
    for (auto& b : bio_) {
        const unsigned p = b->graphPeriod();
        if (p == 0 || B % p != 0) {
            graphNote = "BioSolver " + b->name + " cannot be recorded with blocks of " + std::to_string(B);
            graphSteps_ = 0;
            return false;
        }
    }
    if (forceSolver_.dirty_) {
        graphNote = "host force table edited";
        graphSteps_ = 0;
        return false;
    }
    int stale = 1;
    check(lbib_graph_launch(ctx_, it0, &stale), "graph launch");
    if (!stale) {
        graphLaunches++;
        pending_ += B;
        params_.currentIteration = it0 + B;
        times.steps += B;
        mirrorValid_ = false;
        return true;
    }
    // record the block; the kernels only run when the graph is launched
    const unsigned pend0 = pending_;
    const long steps0 = times.steps;
    check(lbib_capture_begin(ctx_), "capture");
    bool failed = false;
    try {
        for (unsigned j = 0; j < B; j++) enqueueStep(it0 + j);
    } catch (const std::exception& e) {
        failed = true;
        graphNote = std::string("recording failed: ") + e.what();
    }
    int ok = 0;
    check(lbib_capture_end(ctx_, &ok), "capture");
    if (!failed && ok) {
        graphCaptures++;
        check(lbib_graph_launch(ctx_, it0, &stale), "graph launch");
    }
    if (failed || !ok || stale) {
        // nothing ran: back to the launch state of it0
        check(lbib_rewind(ctx_, it0), "rewind");
        pending_ = pend0;
        times.steps = steps0;
        params_.currentIteration = it0;
        forceSolver_.dirty_ = false;
        if (!failed && !ok) graphNote = "the recorded block could not be instantiated";
        if (failed || !ok) graphSteps_ = 0;
        return false;
    }
    graphLaunches++;
    pending_ = pend0 + B;
    params_.currentIteration = it0 + B;
    mirrorValid_ = false;
    return true;
}

void SimulationRunner::run(unsigned steps) {
    const unsigned end = params_.currentIteration + steps;
    const unsigned window = std::max(1u, runAhead);
    while (params_.currentIteration < end) {
        try {
            const unsigned last = std::min(end, params_.currentIteration + window);
            while (params_.currentIteration < last) {
                const unsigned it = params_.currentIteration;
                const unsigned B = graphSteps_;
                if (B && it % B == 1 && it + B <= end && runGraphBlock(it)) continue;
                enqueueStep(it);
            }
            settle();
        } catch (const RunAheadStop& s) {
            resolveStop(s.it);
        }
    }
    if (gh_) Parameters.currentIteration = params_.currentIteration;
}

// SimulationRunner::runSimulation (SimulationRunner.cpp:69-110): Parameters'
// Iterations steps; after step t (1, 2, ...) the reports are written when t is
// a multiple of the report step.  On failure the reports are written for t,
// the state is dumped to output/CRASH_*_<t>.txt and the message printed.
void SimulationRunner::runSimulation() {
    const unsigned start = params_.currentIteration;
    const bool reports = report_ && report_->hasReporters();
    const unsigned rs = std::max(1u, params_.reportSteps);
    try {
        unsigned t = 0;
        while (t < params_.iterations) {
            // run to the next report (or the end) without reading the status
            unsigned next = params_.iterations;
            if (reports) next = std::min(next, (t / rs + 1) * rs);
            run(next - t);
            t = next;
            if (reports && t % rs == 0) {
                settle();
                syncMirror();
                report_->writeReport(t);
            }
        }
    } catch (const std::exception& exp) {
        const unsigned t = params_.currentIteration - start + 1;
        if (report_) {
            try {
                report_->writeCrashReport(t);
            } catch (const std::exception&) {
            }
        }
        const std::string ts = std::to_string(t);
        writeSimulation("output/CRASH_forceOut_" + ts + ".txt", "output/CRASH_geoOut_" + ts + ".txt",
                        "output/CRASH_parameterOut_" + ts + ".txt", "output/CRASH_solverOut_" + ts + ".txt");
        std::cout << exp.what() << std::endl;
    }
}

// vtkFluidReporter.cpp:44-108: density (getRho, slot order of the stored
// state) and velocity (z = 0) at every coarsening-th site
void SimulationRunner::writeVtkFluid(const std::string& prefix, unsigned time, int coarse) {
    if (coarse < 1) LBIB_FAIL("coarsening factor must be positive");
    settle();
    const int nx = (int)params_.sizeX, ny = (int)params_.sizeY;
    const size_t n = (size_t)nx * ny;
    std::vector<double> f(9 * n), ux(n), uy(n);
    check(lbib_get_populations(ctx_, f.data()), "populations");
    check(lbib_get_velocity(ctx_, ux.data(), uy.data()), "velocity");
    std::vector<double> rho, vel;
    // dimensions nx/coarse x ny/coarse, sites (jx coarse, jy coarse)
    for (int jy = 0; jy < ny / coarse; jy++)
        for (int jx = 0; jx < nx / coarse; jx++) {
            const size_t i = (size_t)jy * coarse * nx + (size_t)jx * coarse;
            double r = 0.0;
            for (int q = 0; q < 9; q++) r += f[(size_t)kSlotOfD2Q9[q] * n + i];
            rho.push_back(r);
            vel.push_back(ux[i]);
            vel.push_back(uy[i]);
            vel.push_back(0.0);
        }
    vtkxml::write_image(prefix, time, nx, ny, coarse, {{"density", {1, rho}}, {"velocity", {3, vel}}});
}

// vtkCDEReporter.cpp:44-100: getC of every CDE solver, one array per name
void SimulationRunner::writeVtkCDE(const std::string& prefix, unsigned time, int coarse) {
    if (coarse < 1) LBIB_FAIL("coarsening factor must be positive");
    settle();
    const int nx = (int)params_.sizeX, ny = (int)params_.sizeY;
    const size_t n = (size_t)nx * ny;
    std::vector<std::pair<std::string, std::pair<int, std::vector<double>>>> fields;
    std::vector<double> g(5 * n);
    for (size_t k = 0; k < speciesNames_.size(); k++) {
        check(lbib_get_species(ctx_, (int)k, g.data()), "species");
        std::vector<double> c;
        for (int jy = 0; jy < ny / coarse; jy++)
            for (int jx = 0; jx < nx / coarse; jx++) {
                const size_t i = (size_t)jy * coarse * nx + (size_t)jx * coarse;
                double s = 0.0;
                for (int q = 0; q < 5; q++) s += g[(size_t)kSlotOfD2Q5[q] * n + i];
                c.push_back(s);
            }
        fields.push_back({speciesNames_[k], {1, c}});
    }
    vtkxml::write_image(prefix, time, nx, ny, coarse, fields);
}

void SimulationRunner::writeVtkCells(const std::string& prefix, unsigned time) {
    syncMirror();
    vtkxml::write_cells(prefix, time, geo_, geo_.cellTypeTrackerMap());
}

std::map<unsigned, double> SimulationRunner::computeAccumulatedDomainConcentrations(const std::string& name) {
    const int k = speciesIndex(name);
    if (k < 0) LBIB_FAIL("The CDE Solver with the name " + name + " does not exists.");
    settle();
    const size_t n = (size_t)params_.sizeX * params_.sizeY;
    std::vector<uint32_t> dom(n);
    check(lbib_get_domains(ctx_, dom.data()), "domains");
    uint32_t maxd = 0;
    std::vector<char> present;
    for (size_t i = 0; i < n; i++) {
        if (i % params_.sizeX == 0) continue;   // the reference loops x from 1
        maxd = std::max(maxd, dom[i]);
    }
    present.assign(maxd + 1, 0);
    for (size_t i = 0; i < n; i++)
        if (i % params_.sizeX != 0) present[dom[i]] = 1;
    std::vector<double> acc(maxd + 1);
    check(lbib_domain_concentrations(ctx_, k, (int)maxd + 1, acc.data()), "domain concentrations");
    std::map<unsigned, double> out;
    for (uint32_t d = 0; d <= maxd; d++)
        if (present[d]) out[d] = acc[d];
    return out;
}

std::map<unsigned, double> SimulationRunner::computeAreas() {
    syncMirror();
    std::map<unsigned, int> present;
    unsigned maxdom = 0;
    for (auto& c : geo_.getConnections()) { present[c->domain] = 1; maxdom = std::max(maxdom, c->domain); }
    auto& ctm = geo_.cellTypeTrackerMap();
    for (auto& kv : ctm) maxdom = std::max(maxdom, kv.first);
    std::vector<double> a(maxdom + 1);
    check(lbib_compute_areas(ctx_, a.data()), "computeAreas");
    std::map<unsigned, double> out;
    for (auto& kv : present) out[kv.first] = a[kv.first];
    return out;
}

}  // namespace LbmLib
