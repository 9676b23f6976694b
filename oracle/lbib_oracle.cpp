// lbib_oracle.cpp -- CPU ORACLE (test infrastructure only).
//
// A single-threaded (optionally OpenMP on the race-free lattice loops) C++
// restatement of LBIBCell's per-timestep lattice path, written to follow the
// reference's evaluation order operation by operation so that the FP64 device
// path can be compared bit-for-bit.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library.  The product never does.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off; the reference's release
// flags -O3 -march=native let GCC contract a*b+c into FMA on FMA hosts, which
// changes the last bit of several expressions, including libstdc++'s
// uniform_real_distribution; CMakeLists.txt:49).
//
// Citations are relative to the reference repository root; "LbmLib/" means
// libs/LbmLib/.
//
// State convention.  Like the reference, every site stores its populations in
// the "post-collision, locally swapped" slot order between steps
// (LbmLib/src/solver/FluidSolver/FluidSolver.cpp:259-264,
//  LbmLib/src/solver/CDESolver/CDESolverD2Q5.cpp:151-154), and streaming is the
// in-place pairwise swap of FluidSolver.cpp:185-194 / CDESolverD2Q5.cpp:141-149.
// Accessors convert to natural order (slot d = post-collision f_d).
//
// Parity status (see the tests named in brackets):
//  * D2Q9 collide/stream, Luo force  -- pinned: equilibrium fixed point, mass and
//    momentum balance, shear-wave decay, textbook BGK + Luo step with pull
//    streaming [tests/test_oracle_lattice.py]
//  * Guo force, as written and fixed -- pinned: Guo, Zheng & Shi's source term
//    with the stored velocity, NE coefficient difference [test_oracle_pins.py]
//  * D2Q5 species                     -- pinned: mass balance, Gaussian spreading
//    with D=(tau-1/2)/3 against the reference's DiracD2Q5 closed form; SIGNAL
//    source/decay balances; LINEAR species as the (I+K) matrix map on a
//    uniform field; the Schnakenberg explicit-map recurrence
//  * no-flux D2Q5 boundary            -- pinned: hand-derived incoming values,
//    first species only and the swap into every listed species [pins]
//  * CDE re-initialisation            -- pinned: hand-derived means (direct,
//    diagonal, SIGNAL fallback, NaN of the plain solver) then collide [pins]
//  * fastCos / delta                 -- pinned: table vs (1+cos)/4 within the
//    approximation error, partition of unity, evenness defect
//  * spread / interpolate             -- pinned: adjointness <S F,u> = <F,I u>
//  * cut links + scanline             -- pinned: brute-force point-in-polygon on
//    small grids, count inside a circle ~ pi R^2; closest-wins tie goes to the
//    earlier connection; boundary-node velocity linear along the segment [pins]
//  * shoelace area                    -- pinned: regular n-gon closed form;
//    parity against boost::geometry::area itself is UNPINNED (Boost absent)
//  * tension force type 6             -- pinned: regular n-gon net inward force
//  * mt19937(1) + uniform_real        -- pinned: [rand.predef] KAT 4123659995 and
//    an independent MT19937; node and connection draw order [pins]
//  * remesh / coarsen                 -- pinned: hand-built rings with known ids,
//    midpoints, removal sequence and connection order [pins]

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---- constants: LbmLib/include/Constants.hpp:31-64, Direction.hpp:34-44 ----
enum Dir { T = 0, E = 1, N = 2, W = 3, S = 4, NE = 5, NW = 6, SW = 7, SE = 8 };
static const int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
// Direction.cpp:27-61
static const int INV[9] = {T, W, S, E, N, SW, SE, NE, NW};
static const double CS = 1 / std::sqrt(3);   // Constants.hpp:33 (compared to |u|^2)
static const double EPSILON = 10E-10;        // Constants.hpp:34
static const double PERTURBATION = 10E-7;    // Constants.hpp:35
static const double MAXLENGTH = 0.5;         // Constants.hpp:49
// FluidSolver.cpp:39-47
static const double W0 = 4.0 / 9.0;
static const double W1 = 1.0 / 9.0;
static const double W2 = 1.0 / 36.0;

struct Fail : std::runtime_error {
    explicit Fail(const std::string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------- species --
// Reaction kinds the registry serves (CDESolverFactory name -> spec).
enum SpeciesKind {
    SP_PLAIN = 0,    // CDESolverD2Q5 (CDESolverD2Q5.cpp:37-228)
    SP_SIGNAL = 1,   // tutorial_01_CDESolverD2Q5_SIGNAL (…_SIGNAL.cpp:39-250)
    SP_LINEAR = 2,   // linear reaction R_s = sum_j K[s][j] C_j (+source by domain)
    SP_SCHNAKENBERG_R = 3,   // tutorial_02_CDESolverD2Q5_R (…_R.cpp:38-42, 94-140)
    SP_SCHNAKENBERG_L = 4,   // tutorial_02_CDESolverD2Q5_L (…_L.cpp:38-42, 94-141)
};

struct Species {
    int kind = SP_PLAIN;
    double tau = 1.0;
    // SIGNAL parameters (…_SIGNAL.cpp:39-43): production in domain `prod_domain`
    // (the reference hard-codes 1) while it < switchoff, decay elsewhere;
    // init value inside domains != 0.
    double production = 0.0001, decay = 0.00001, init_inside = 1.0;
    unsigned switchoff = 5000;
    int prod_domain = 1;      // -1: every domain != 0 produces
    // LINEAR parameters: row of the coupling matrix and a source inside cells.
    std::vector<double> krow;  // length = number of species
    double source_inside = 0.0;
    // Schnakenberg pair: gamma, a (R) or b (L), deltaT, index of the partner
    double gamma = 100.0, ab = 0.1, deltaT = 1.0e-4;
    int partner = 0;
    std::string name;
};

// ---------------------------------------------------------------- geometry --
struct Conn;
struct Marker {                       // GeometryNode (GeometryNode.cpp:98-117)
    unsigned id;
    double x, y;                      // LagrangianPoint::position_
    double vx = 0, vy = 0;            // LagrangianPoint::velocity_
    double fx = 0, fy = 0;            // GeometryNode::force_
    double ox, oy;                    // oldPosition_  (ctor: x-5, y-5)
    std::array<long, 16> stencil;     // neighbourPhysicalNodes_ (site index)
    Conn* c0 = nullptr;               // incoming connection (getConnection<0>)
    Conn* c1 = nullptr;               // outgoing connection (getConnection<1>)
};

struct Conn {                         // Connection.hpp:47-155
    Marker* a;                        // points_.first
    Marker* b;                        // points_.second
    unsigned domain;
    std::vector<int> bc_species;      // NoFluxD2Q5 descriptor, species indices
};

struct BNode {                        // BoundaryNode.cpp:35-69
    double x, y;
    unsigned domain;
    double vx = 0, vy = 0;
    long site = -1;                   // physicalNeighbour_
    int dir = 0;                      // directionToPhysicalNeighbour_
    std::vector<int> bc_species;      // boundary solvers built from the descriptor (BoundaryNode.cpp:43-46)
    double dist = 0;                  // BC scratch (BoundarySolverNoFluxD2Q5 distribution_)
    bool alive = true;
    int conn = -1;                    // index of the generating connection in Sim::conns
};

struct ForceStruct {                  // ForceStructs.hpp:49-588
    int type;
    unsigned id1, id2;
    double a, b;                      // k,l0 | f | x,y,k,l0 packed (see calc)
    double c = 0, d = 0;
};

struct Sim {
    int nx, ny;
    double tauf;
    int force_model = 0;   // 0 Luo (FluidSolver.hpp:41), 1 Guo as written, 2 Guo fixed
    unsigned it = 0;       // Parameters.currentIteration_
    // lattice (AoS per site, swapped slot order between steps)
    std::vector<std::array<double, 9>> f;
    std::vector<double> ux, uy, frx, fry, madd;
    std::vector<unsigned> dom, ctype;
    std::vector<std::array<long, 5>> bn;           // boundaryNodes_[T..S]
    std::vector<Species> sp;
    std::vector<std::vector<std::array<double, 5>>> g;   // [species][site]
    // geometry
    std::map<unsigned, std::unique_ptr<Marker>> markers;
    std::vector<std::shared_ptr<Conn>> conns;
    std::vector<BNode> bnodes;
    std::map<unsigned, unsigned> celltype_map;
    // forces
    std::map<unsigned, std::vector<ForceStruct>> forcemap;
    // RNG streams (GeometryNode.cpp:155-158, Connection.cpp:334-337)
    std::mt19937 gen_node{1}, gen_conn{1};
    std::uniform_real_distribution<> dis_node{-PERTURBATION, PERTURBATION};
    std::uniform_real_distribution<> dis_conn{-PERTURBATION, PERTURBATION};
    long n_perturb_node = 0, n_perturb_conn = 0;
    // plugins
    bool box_outlet = false;
    struct Bio {
        int kind; double p[4]; unsigned freq;
        std::map<unsigned, double> target, integ;   // area regulators
        bool inited = false;                        // differentiation's call_once
        std::mt19937 gen{1};                        // random-direction division
    };
    std::vector<Bio> bio;
    int threads = 1;
    std::vector<unsigned> reinit_log;  // sites reinitialised during the last step
    // seconds per stage: forces + spread, boundary + streaming, collide,
    // interpolation + move + cut links + domain ids, remesh + coarsen,
    // mass solver + BioSolvers
    double stage_s[6] = {0, 0, 0, 0, 0, 0};
    std::string last_error;

    long idx(long x, long y) const { return y * nx + x; }
};

// ------------------------------------------------------------- fastCos, δ --
// GeometryNode.cpp:47-94
static const double halfPi = M_PI / 2.0;
static const double doublePi = 2.0 * M_PI;
static inline double fastCos(const double x) {
    double c;
    double t = x + halfPi;
    if (t > M_PI) t -= doublePi;
    if (t < .00) {
        c = 1.27323954 * t + 0.405284735 * t * t;
        if (c < 0.0) c = 0.225 * (c * -c - c) + c;
        else         c = 0.225 * (c * c - c) + c;
    } else {
        c = 1.27323954 * t - 0.405284735 * t * t;
        if (c < 0.0) c = 0.225 * (c * -c - c) + c;
        else         c = 0.225 * (c * c - c) + c;
    }
    return c;
}
static inline double delta(const double dx, const double dy) {
    return ((1.0 + fastCos(halfPi * dx)) * (1.0 + fastCos(halfPi * dy))) / 16.0;
}

// ------------------------------------------------------------ perturbation --
// GeometryNode::perturbatePosition, GeometryNode.cpp:153-214
static void perturbatePosition(Sim& s, Marker& m) {
    if (std::abs(static_cast<int>(m.x) - m.x) <= EPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * EPSILON) r = s.dis_node(s.gen_node);
        m.x = m.x + r; s.n_perturb_node++;
    }
    if (std::abs(static_cast<int>(m.y) - m.y) <= EPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * EPSILON) r = s.dis_node(s.gen_node);
        m.y = m.y + r; s.n_perturb_node++;
    }
    // setPos(getXPos() + dis(gen), getYPos() + dis(gen)): GCC evaluates the call
    // arguments right to left, so the first draw goes to y.
    if (std::abs(m.y - std::floor(m.y) - (m.x - std::floor(m.x))) <= EPSILON) {
        double ry = s.dis_node(s.gen_node);
        double rx = s.dis_node(s.gen_node);
        m.x = m.x + rx; m.y = m.y + ry; s.n_perturb_node++;
    }
    if (std::abs(std::ceil(m.y) - m.y - (m.x - std::floor(m.x))) <= EPSILON) {
        double ry = s.dis_node(s.gen_node);
        double rx = s.dis_node(s.gen_node);
        m.x = m.x + rx; m.y = m.y + ry; s.n_perturb_node++;
    }
}
static void setPos(Sim& s, Marker& m, double x, double y) {   // GeometryNode.cpp:143-151
    m.x = x; m.y = y;
    perturbatePosition(s, m);
}
// Connection::perturbConnection, Connection.cpp:332-387
static void perturbConnection(Sim& s, Conn& c) {
    if (std::abs(c.a->x - c.b->x) < EPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * EPSILON) r = s.dis_conn(s.gen_conn);
        setPos(s, *c.a, c.a->x + r, c.a->y); s.n_perturb_conn++;
    }
    if (std::abs(c.a->y - c.b->y) < EPSILON) {
        double r = 0.0;
        while (std::abs(r) < 5.0 * EPSILON) r = s.dis_conn(s.gen_conn);
        setPos(s, *c.a, c.a->x, c.a->y + r); s.n_perturb_conn++;
    }
}
// GeometryHandler::perturbConnections, GeometryHandler.cpp:326-356
static void perturbConnections(Sim& s) {
    std::map<unsigned, Conn*> first;
    for (auto& c : s.conns)
        if (!first.count(c->domain)) first[c->domain] = c.get();
    for (auto& kv : first) {
        Conn* start = kv.second;
        Conn* t = start;
        do {
            perturbConnection(s, *t);
            t = t->b->c1;
        } while (t != start);
    }
}

// ----------------------------------------------------------------- stencil --
// GeometryNode::isUpdateNeeded (GeometryNode.cpp:256-261) with
// Rectangle::pointWithinBounds (UtilLib/src/geometry/Rectangle.cpp:63-71)
static bool isUpdateNeeded(const Marker& m) {
    const double north = std::ceil(m.oy), south = std::floor(m.oy);
    const double west = std::floor(m.ox), east = std::ceil(m.ox);
    const bool within = (m.x >= west) && (m.x < east) && (m.y <= north) && (m.y > south);
    return !within;
}
// GeometryHandler::connectGeometryNodesToPhysicalNodes, GeometryHandler.cpp:608-644
static void connectMarker(Sim& s, Marker& m) {
    if (!isUpdateNeeded(m)) return;
    const double minX = m.x - 2.0, maxX = m.x + 2.0, minY = m.y - 2.0, maxY = m.y + 2.0;
    if (minX < 0) throw Fail("geometry point moved too close to the west boundary");
    if (minY < 0) throw Fail("geometry point moved too close to the south boundary");
    if (maxX > s.nx) throw Fail("geometry point moved too close to the east boundary");
    if (maxY > s.ny) throw Fail("geometry point moved too close to the north boundary");
    for (int x = (int)std::ceil(minX); x <= std::floor(maxX); x++)
        for (int y = (int)std::ceil(minY); y <= std::floor(maxY); y++)
            m.stencil[(x * 4 + y) % 16] = s.idx(x, y);
}
static void connectAll(Sim& s) {
    for (auto& kv : s.markers) connectMarker(s, *kv.second);
}

// --------------------------------------------------------------- cut links --
static double sqd(double x1, double y1, double x2, double y2) {   // EulerianPoint.hpp:88-96
    return (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2);
}
// Connection::calculateVelocity, Connection.cpp:146-168
static void bnVelocity(const Conn& c, BNode& b) {
    const double l = std::sqrt(sqd(c.a->x, c.a->y, b.x, b.y));
    const double L = std::sqrt(sqd(c.b->x, c.b->y, c.a->x, c.a->y));
    b.vx = (c.b->vx - c.a->vx) * l / L + c.a->vx;
    b.vy = (c.b->vy - c.a->vy) * l / L + c.a->vy;
}
// Connection::insertBoundaryNode, Connection.cpp:181-222
static void insertBN(Sim& s, const Conn& c, BNode nb, long site, int d) {
    const int slot = INV[d];
    long old = s.bn[site][slot];
    const double px = (double)(site % s.nx), py = (double)(site / s.nx);
    if (old < 0 || sqd(px, py, nb.x, nb.y) < sqd(px, py, s.bnodes[old].x, s.bnodes[old].y)) {
        bnVelocity(c, nb);
        nb.site = site; nb.dir = d; nb.bc_species = c.bc_species;
        if (old >= 0) s.bnodes[old].alive = false;
        s.bnodes.push_back(nb);
        s.bn[site][slot] = (long)s.bnodes.size() - 1;   // setPhysicalNeighbours
    }
}
// Connection::generateBoundaryNodes, Connection.cpp:224-324
static void generateBN(Sim& s, const Conn& c, int ci) {
    const double x1 = c.a->x, y1 = c.a->y, x2 = c.b->x, y2 = c.b->y;
    const double maxX = std::max(x1, x2), maxY = std::max(y1, y2);
    const double minX = std::min(x1, x2), minY = std::min(y1, y2);
    const double m = (y2 - y1) / (x2 - x1);
    const double n = y1 - (y2 - y1) / (x2 - x1) * x1;
    if (!std::isfinite(m))
        throw Fail("This can only happen if something went wrong with the perturbation of the connection.");
    for (int i = (int)std::ceil(minX); i <= std::floor(maxX); i++) {
        const double y = m * i + n;
        BNode bN{(double)i, y, 0}, bS{(double)i, y, 0};
        bN.conn = bS.conn = ci;
        if (x1 < x2) { bN.domain = c.domain; bS.domain = 0; }
        else         { bN.domain = 0; bS.domain = c.domain; }
        const long sS = s.idx(i, (long)std::floor(y));
        const long sN = s.idx(i, (long)std::ceil(y));
        insertBN(s, c, bS, sS, S);
        insertBN(s, c, bN, sN, N);
    }
    for (int i = (int)std::ceil(minY); i <= std::floor(maxY); i++) {
        const double x = (i - n) / m;
        BNode bW{x, (double)i, 0}, bE{x, (double)i, 0};
        bW.conn = bE.conn = ci;
        if (y1 > y2) { bW.domain = 0; bE.domain = c.domain; }
        else         { bW.domain = c.domain; bE.domain = 0; }
        const long sW = s.idx((long)std::floor(x), i);
        const long sE = s.idx((long)std::ceil(x), i);
        insertBN(s, c, bW, sW, W);
        insertBN(s, c, bE, sE, E);
    }
}
static void regenerateBoundaryNodes(Sim& s) {
    for (auto& a : s.bn) a = {-1, -1, -1, -1, -1};
    s.bnodes.clear();
    for (size_t k = 0; k < s.conns.size(); k++)   // GeometryHandler.cpp:646-659 (serial)
        generateBN(s, *s.conns[k], (int)k);
}

// ------------------------------------------------------------ fluid solver --
static inline double rhoOfSlots(const std::array<double, 9>& d) {   // FluidSolver.cpp:248-251
    double r = 0.0;
    for (int q = 0; q < 9; q++) r += d[q];
    return r;
}
static inline void localSwap9(std::array<double, 9>& d) {          // FluidSolver.cpp:259-264
    std::swap(d[E], d[W]); std::swap(d[NE], d[SW]); std::swap(d[N], d[S]); std::swap(d[NW], d[SE]);
}
static inline void localSwap5(std::array<double, 5>& d) {          // CDESolverD2Q5.cpp:151-154
    std::swap(d[E], d[W]); std::swap(d[N], d[S]);
}
// FluidSolver::collide (FluidSolver.cpp:103-183) with Luo1993ForceModel
// (Luo1993ForceModel.cpp:29-70) or GuoZhengShi2002ForceModel (…:29-78).
static void fluidCollide(Sim& s, long i) {
    std::array<double, 9>& d = s.f[i];
    const double rho0 = rhoOfSlots(d);
    double rho = rho0;
    if (std::isfinite(s.madd[i])) rho += s.madd[i];
    const double Fx = s.frx[i], Fy = s.fry[i];
    const bool guo = s.force_model != 0;
    auto velo = [&](double& ux, double& uy) {
        const double rhoI = 1.0 / rho;
        if (!guo) {
            ux = (d[E] + d[NE] + d[SE] - (d[NW] + d[W] + d[SW])) * rhoI;
            uy = (d[NE] + d[N] + d[NW] - (d[SW] + d[S] + d[SE])) * rhoI;
        } else {
            ux = (d[E] + d[NE] + d[SE] - (d[NW] + d[W] + d[SW]) + 0.5 * Fx) * rhoI;
            uy = (d[NE] + d[N] + d[NW] - (d[SW] + d[S] + d[SE]) + 0.5 * Fy) * rhoI;
        }
    };
    double ux, uy;
    velo(ux, uy);
    const double ux2 = ux * ux, uy2 = uy * uy;
    const double lastTerm = 3.0 / 2.0 * (ux2 + uy2);
    const double tauI = 1.0 / s.tauf;
    const double W3 = 9.0 / 2.0;
    const double W1rho = W1 * rho, W2rho = W2 * rho;
    double fc[9];
    if (!guo) {
        const double W1rho3 = rho0 * 3.0 / 9.0, W2rho3 = rho0 * 3.0 / 36.0;
        fc[T] = 0.0;
        fc[E] = W1rho3 * Fx;  fc[N] = W1rho3 * Fy;
        fc[W] = -W1rho3 * Fx; fc[S] = -W1rho3 * Fy;
        fc[NE] = W2rho3 * (Fx + Fy);  fc[NW] = W2rho3 * (-Fx + Fy);
        fc[SW] = W2rho3 * (-Fx - Fy); fc[SE] = W2rho3 * (Fx - Fy);
    } else {
        // Guo uses the *stored* velocity_ of the previous collide (FluidSolver.cpp:141-145)
        const double gx = s.ux[i], gy = s.uy[i], tau = s.tauf;
        const double G0 = (1.0 - 0.5 / tau) * 4.0 / 9.0;
        const double G1 = (1.0 - 0.5 / tau) / 9.0;
        const double G2 = (1.0 - 0.5 / tau) / 36.0;
        fc[T] = G0 * (-3.0 * gx * Fx - 3.0 * gy * Fy);
        fc[E] = G1 * ((6.0 * gx + 3.0) * Fx - 3.0 * gy * Fy);
        fc[N] = G1 * (-3.0 * gx * Fx + (6.0 * gy + 3.0) * Fy);
        fc[W] = G1 * ((6.0 * gx - 3.0) * Fx - 3.0 * gy * Fy);
        fc[S] = G1 * (-3.0 * gx * Fx + (6.0 * gy - 3.0) * Fy);
        // GuoZhengShi2002ForceModel.cpp:51 carries (9ux - 3uy + 3) in the NE y-term;
        // force_model 2 uses Guo et al.'s (9ux + 6uy + 3).
        if (s.force_model == 1)
            fc[NE] = G2 * ((6.0 * gx + 9.0 * gy + 3.0) * Fx + (9.0 * gx - 3.0 * gy + 3.0) * Fy);
        else
            fc[NE] = G2 * ((6.0 * gx + 9.0 * gy + 3.0) * Fx + (9.0 * gx + 6.0 * gy + 3.0) * Fy);
        fc[NW] = G2 * ((6.0 * gx - 9.0 * gy - 3.0) * Fx + (-9.0 * gx + 6.0 * gy + 3.0) * Fy);
        fc[SW] = G2 * ((6.0 * gx + 9.0 * gy - 3.0) * Fx + (9.0 * gx + 6.0 * gy - 3.0) * Fy);
        fc[SE] = G2 * ((6.0 * gx - 9.0 * gy + 3.0) * Fx + (-9.0 * gx + 6.0 * gy - 3.0) * Fy);
    }
    double tmp[9];
    tmp[T] = W0 * rho * (1.0 - lastTerm) * tauI + fc[T];
    tmp[E] = W1rho * (1.0 + 3.0 * ux + W3 * ux2 - lastTerm) * tauI + fc[E];
    tmp[NE] = W2rho * (1.0 + 3.0 * (ux + uy) + W3 * (ux + uy) * (ux + uy) - lastTerm) * tauI + fc[NE];
    tmp[N] = W1rho * (1.0 + 3.0 * uy + W3 * uy2 - lastTerm) * tauI + fc[N];
    tmp[NW] = W2rho * (1.0 + 3.0 * (-ux + uy) + W3 * (-ux + uy) * (-ux + uy) - lastTerm) * tauI + fc[NW];
    tmp[W] = W1rho * (1.0 + 3.0 * (-ux) + W3 * ux2 - lastTerm) * tauI + fc[W];
    tmp[SW] = W2rho * (1.0 + 3.0 * (-ux - uy) + W3 * (-ux - uy) * (-ux - uy) - lastTerm) * tauI + fc[SW];
    tmp[S] = W1rho * (1.0 + 3.0 * (-uy) + W3 * uy2 - lastTerm) * tauI + fc[S];
    tmp[SE] = W2rho * (1.0 + 3.0 * (ux - uy) + W3 * (ux - uy) * (ux - uy) - lastTerm) * tauI + fc[SE];
    for (int q = 0; q < 9; q++) {
        const double td = d[q];
        d[q] = td - td * tauI + tmp[q];
    }
    velo(ux, uy);
    s.ux[i] = ux; s.uy[i] = uy;
    s.madd[i] = 0.0;
    localSwap9(d);
}
static void fluidInit(Sim& s, long i) {                 // FluidSolver.cpp:50-60
    s.f[i] = {1.0, 0, 0, 0, 0, 0, 0, 0, 0};
    fluidCollide(s, i);
}
// FluidSolver::initWithVelocity, FluidSolver.cpp:62-101
static void fluidInitVelocity(Sim& s, long i, double u, double v) {
    const double rho = 1.0, u2 = u * u, v2 = v * v;
    const double W1rho = W1 * rho, W2rho = W2 * rho, W3 = 9.0 / 2.0;
    const double lastTerm = 3.0 / 2.0 * (u2 + v2);
    std::array<double, 9>& d = s.f[i];
    d[T] = W0 * rho * (1.0 - lastTerm);
    d[E] = W1rho * (1.0 + 3.0 * u + W3 * u2 - lastTerm);
    d[NE] = W2rho * (1.0 + 3.0 * (u + v) + W3 * (u + v) * (u + v) - lastTerm);
    d[N] = W1rho * (1.0 + 3.0 * v + W3 * v2 - lastTerm);
    d[NW] = W2rho * (1.0 + 3.0 * (-u + v) + W3 * (-u + v) * (-u + v) - lastTerm);
    d[W] = W1rho * (1.0 + 3.0 * (-u) + W3 * u2 - lastTerm);
    d[SW] = W2rho * (1.0 + 3.0 * (-u - v) + W3 * (-u - v) * (-u - v) - lastTerm);
    d[S] = W1rho * (1.0 + 3.0 * (-v) + W3 * v2 - lastTerm);
    d[SE] = W2rho * (1.0 + 3.0 * (u - v) + W3 * (u - v) * (u - v) - lastTerm);
    s.ux[i] = u; s.uy[i] = v;
    s.madd[i] = 0.0;
    localSwap9(d);
}

// --------------------------------------------------------------- species ---
static inline double cOfSlots(const std::array<double, 5>& d) {    // CDESolverD2Q5.cpp:71-73
    double c = 0.0;
    for (int q = 0; q < 5; q++) c += d[q];
    return c;
}
// Reaction term of species k at site i.  `cj` holds every species'
// concentration at the site as seen by this collide (cj[k] == C).  The linear
// spec is this framework's multi-species registry entry (the reference ships
// per-pair C++ plugins such as SchnakenbergD2Q5u/v); it is evaluated Jacobi-style
// on the pre-collision concentrations of all species.
static double reaction(const Sim& s, int k, long i, double C, const double* cj) {
    const Species& p = s.sp[k];
    if (p.kind == SP_SIGNAL) {      // …_SIGNAL.cpp:92-100
        const bool prod = (p.prod_domain < 0) ? (s.dom[i] != 0) : (s.dom[i] == (unsigned)p.prod_domain);
        if (prod && (s.it < p.switchoff)) return p.production;
        return -p.decay * C;
    }
    if (p.kind == SP_LINEAR) {
        double r = 0.0;
        for (size_t j = 0; j < p.krow.size(); j++) r += p.krow[j] * cj[j];
        if (s.dom[i] != 0) r += p.source_inside;
        return r;
    }
    if (p.kind == SP_SCHNAKENBERG_R || p.kind == SP_SCHNAKENBERG_L) {
        // …_R.cpp:94-105 / …_L.cpp:94-106: the partner's getC() on its array as
        // it is stored right now (already collided and swapped at this site
        // when it comes first in the solver order)
        const double Cp = cOfSlots(s.g[p.partner][i]);
        if (s.dom[i] == 0) return 0.0;
        if (p.kind == SP_SCHNAKENBERG_R) {
            const double R = C, L = Cp;
            return p.gamma * (p.ab - R + R * R * L);
        }
        const double R = Cp, L = C;
        return p.gamma * (p.ab - R * R * L);
    }
    return 0.0;
}
// CDESolverD2Q5::collide (CDESolverD2Q5.cpp:75-107) and the SIGNAL variant
// (…_SIGNAL.cpp:102-135).  u is the same site's fluid velocity_ after this
// step's fluid collide (SimulationRunner.cpp:251-258).
static void speciesCollide(Sim& s, int k, long i, const double* cj) {
    std::array<double, 5>& d = s.g[k][i];
    const double C = cOfSlots(d);
    const double ux = s.ux[i], uy = s.uy[i];
    const double w0 = C / 3.0, w1 = C / 6.0;
    const double tauI = 1.0 / s.sp[k].tau;
    double tmp[5];
    tmp[T] = w0;
    tmp[E] = w1 * (1.0 + ux * 3.0);
    tmp[N] = w1 * (1.0 + uy * 3.0);
    tmp[W] = w1 * (1.0 + (-ux) * 3.0);
    tmp[S] = w1 * (1.0 + (-uy) * 3.0);
    if (s.sp[k].kind == SP_PLAIN) {
        for (int q = 0; q < 5; q++) {
            const double td = d[q];
            d[q] = td - td * tauI + tmp[q] * tauI;
        }
    } else {
        const double R = reaction(s, k, i, C, cj);
        const bool sch = s.sp[k].kind == SP_SCHNAKENBERG_R || s.sp[k].kind == SP_SCHNAKENBERG_L;
        // …_R.cpp:128-130: deltaT*reactionTerm / 3.0 and / 6.0
        const double R0 = sch ? s.sp[k].deltaT * R / 3.0 : 1.0 * R / 3.0;
        const double R1 = sch ? s.sp[k].deltaT * R / 6.0 : 1.0 * R / 6.0;
        d[T] = d[T] - d[T] * tauI + tmp[T] * tauI + R0;
        d[E] = d[E] - d[E] * tauI + tmp[E] * tauI + R1;
        d[N] = d[N] - d[N] * tauI + tmp[N] * tauI + R1;
        d[W] = d[W] - d[W] * tauI + tmp[W] * tauI + R1;
        d[S] = d[S] - d[S] * tauI + tmp[S] * tauI + R1;
    }
    localSwap5(d);
}
static void speciesInit(Sim& s, int k, long i) {
    const Species& p = s.sp[k];
    if (p.kind != SP_PLAIN) {       // …_SIGNAL.cpp:46-57, …_R.cpp:44-58 (no collide;
                                    // the R/L noise is the caller's)
        const double v = (s.dom[i] == 0) ? 0.0 : p.init_inside / 5.0;
        s.g[k][i] = {v, v, v, v, v};
    } else {                        // CDESolverD2Q5.cpp:37-40
        s.g[k][i] = {1.0, 0, 0, 0, 0};
        speciesCollide(s, k, i, nullptr);
    }
}
// CDESolverD2Q5::reinitialise (CDESolverD2Q5.cpp:156-228) / SIGNAL variant
// (…_SIGNAL.cpp:184-250).  Diagonal boundary-neighbour reads index past the
// 5-slot array in the reference (undefined behaviour); they are taken as absent.
static void speciesReinit(Sim& s, int k, long i) {
    const long x = i % s.nx, y = i / s.nx;
    auto nb = [&](int d) { return s.idx((x + EX[d] + s.nx) % s.nx, (y + EY[d] + s.ny) % s.ny); };
    double sumC = 0.0;
    int counter = 0;
    const unsigned my = s.dom[i];
    const int dir1[4] = {E, N, W, S};
    for (int d : dir1) {
        const long j = nb(d);
        if (s.bn[i][d] < 0 && my == s.dom[j]) { sumC += cOfSlots(s.g[k][j]); counter++; }
    }
    if (counter == 0) {
        const int dir2[4] = {NE, NW, SW, SE};
        for (int d : dir2) {
            const long j = nb(d);
            if (my == s.dom[j]) { sumC += cOfSlots(s.g[k][j]); counter++; }
        }
    }
    if (counter == 0 && s.sp[k].kind != SP_PLAIN) {
        for (int d : dir1) { sumC += cOfSlots(s.g[k][nb(d)]); counter++; }
    }
    sumC /= static_cast<double>(counter);
    for (int q = 0; q < 5; q++) s.g[k][i][q] = sumC / 5.0;
    double cj[16];
    for (size_t j = 0; j < s.sp.size() && j < 16; j++) cj[j] = cOfSlots(s.g[j][i]);
    speciesCollide(s, k, i, cj);
}

// ------------------------------------------------------- streaming (swap) ---
// FluidSolver::advect (FluidSolver.cpp:185-194), CDESolverD2Q5::advect (:141-149)
static void advect(Sim& s) {
    const int nx = s.nx, ny = s.ny;
#pragma omp parallel for schedule(static) num_threads(s.threads) if (s.threads > 1)
    for (long y = 0; y < ny; y++) {
        for (long x = 0; x < nx; x++) {
            const long i = y * nx + x;
            const long iW = y * nx + (x - 1 + nx) % nx;
            const long yS = (y - 1 + ny) % ny;
            const long iSW = yS * nx + (x - 1 + nx) % nx;
            const long iS = yS * nx + x;
            const long iSE = yS * nx + (x + 1) % nx;
            std::swap(s.f[i][E], s.f[iW][W]);
            std::swap(s.f[i][NE], s.f[iSW][SW]);
            std::swap(s.f[i][N], s.f[iS][S]);
            std::swap(s.f[i][NW], s.f[iSE][SE]);
            for (size_t k = 0; k < s.sp.size(); k++) {
                std::swap(s.g[k][i][E], s.g[k][iW][W]);
                std::swap(s.g[k][i][N], s.g[k][iS][S]);
            }
        }
    }
}

// --------------------------------------------------------- boundary solver --
// BoundarySolverNoFluxD2Q5::preAdvect (BoundarySolverNoFluxD2Q5.cpp:51-113):
// only the first species of the descriptor gets a value (the loop returns).
static void preAdvect(Sim& s) {
    for (auto& b : s.bnodes) {
        if (!b.alive || b.bc_species.empty()) continue;
        const int k = b.bc_species[0];
        const auto& d = s.g[k][b.site];
        const double C = cOfSlots(d);
        const double Cf = C / 6.0;
        const double ux = b.vx, uy = b.vy;
        // calculateEquilibrium(dir) with the site's stored velocity_
        const double pu = s.ux[b.site], pv = s.uy[b.site];
        const double w1 = C / 6.0;
        double feq = 0, eq = 0;
        switch (b.dir) {
            case E: feq = Cf * (1 + 3 * ux); eq = w1 * (1.0 + pu * 3.0); break;
            case N: feq = Cf * (1 + 3 * uy); eq = w1 * (1.0 + pv * 3.0); break;
            case W: feq = Cf * (1 - 3 * ux); eq = w1 * (1.0 + (-pu) * 3.0); break;
            case S: feq = Cf * (1 - 3 * uy); eq = w1 * (1.0 + (-pv) * 3.0); break;
        }
        const double fneq = d[INV[b.dir]] - eq;
        b.dist = feq + fneq;
    }
}
// BoundarySolverNoFluxD2Q5::postAdvect (:37-47): swaps into every listed species.
static void postAdvect(Sim& s) {
    for (auto& b : s.bnodes) {
        if (!b.alive || b.bc_species.empty()) continue;
        for (int k : b.bc_species) std::swap(b.dist, s.g[k][b.site][b.dir]);
    }
}

// ----------------------------------------------------------------- forces --
// ForceSolver::calculateForce (ForceSolver.cpp:42-85); ForceStructs.hpp:49-588
static Marker& node(Sim& s, unsigned id) {
    auto it = s.markers.find(id);
    if (it == s.markers.end()) throw Fail("Could not compute and set the forces.");
    return *it->second;
}
static void calculateForce(Sim& s) {
    for (auto& kv : s.markers) { kv.second->fx = 0.0; kv.second->fy = 0.0; }
    for (auto& kv : s.forcemap) {
        for (auto& fs : kv.second) {
            Marker& m1 = node(s, fs.id1);
            switch (fs.type) {
                case 0: case 7: {      // spring k(|d|-l0) d/|d|
                    Marker& m2 = node(s, fs.id2);
                    const double f = fs.a * (std::sqrt(sqd(m2.x, m2.y, m1.x, m1.y)) - fs.b);
                    const double dist = std::sqrt(sqd(m2.x, m2.y, m1.x, m1.y));
                    m1.fx += f * (m2.x - m1.x) / dist;
                    m1.fy += f * (m2.y - m1.y) / dist;
                    break;
                }
                case 5: case 6: {      // constant f d/|d|
                    Marker& m2 = node(s, fs.id2);
                    const double dist = std::sqrt(sqd(m2.x, m2.y, m1.x, m1.y));
                    m1.fx += fs.a * (m2.x - m1.x) / dist;
                    m1.fy += fs.a * (m2.y - m1.y) / dist;
                    break;
                }
                case 1: {              // anchor at (a,b), k=c, l0=d
                    const double dist0 = std::sqrt(sqd(fs.a, fs.b, m1.x, m1.y));
                    const double f = fs.c * (dist0 - fs.d);
                    const double dist = std::sqrt(sqd(fs.a, fs.b, m1.x, m1.y));
                    if (dist != 0.0) { m1.fx += f * (fs.a - m1.x) / dist; m1.fy += f * (fs.b - m1.y) / dist; }
                    break;
                }
                case 2: m1.fx += fs.a; m1.fy += fs.b; break;
                case 3: {              // horizontal slider y=a, k=b, l0=c
                    const double f = fs.b * (std::sqrt(sqd(m1.x, fs.a, m1.x, m1.y)) - fs.c);
                    m1.fx += 0; m1.fy += (m1.y < fs.a) ? f : -f;
                    break;
                }
                case 4: {              // vertical slider x=a, k=b, l0=c
                    const double f = fs.b * (std::sqrt(sqd(fs.a, m1.y, m1.x, m1.y)) - fs.c);
                    m1.fx += (m1.x < fs.a) ? f : -f; m1.fy += 0;
                    break;
                }
            }
        }
    }
}
// SimulationRunner::distributeForce (SimulationRunner.cpp:425-453) ->
// GeometryNode::distributeForce (GeometryNode.cpp:228-235), marker-id order.
static void distributeForce(Sim& s) {
    std::fill(s.frx.begin(), s.frx.end(), 0.0);
    std::fill(s.fry.begin(), s.fry.end(), 0.0);
    for (auto& kv : s.markers) {
        const Marker& m = *kv.second;
        for (int k = 0; k < 16; k++) {
            const long site = m.stencil[k];
            const double dx = (double)(site % s.nx) - m.x, dy = (double)(site / s.nx) - m.y;
            const double dd = delta(dx, dy);
            s.frx[site] += dd * m.fx;
            s.fry[site] += dd * m.fy;
        }
    }
}
// GeometryNode::collectVelocity (GeometryNode.cpp:216-226) + the velocity guard
// of LagrangianPoint::setVelocity (LagrangianPoint.cpp:92-103).
static void collectVelocity(Sim& s) {
    for (auto& kv : s.markers) {
        Marker& m = *kv.second;
        double vx = 0.0, vy = 0.0;
        for (int k = 0; k < 16; k++) {
            const long site = m.stencil[k];
            const double dx = (double)(site % s.nx) - m.x, dy = (double)(site / s.nx) - m.y;
            const double dd = delta(dx, dy);
            vx += dd * s.ux[site];
            vy += dd * s.uy[site];
        }
        m.vx = vx; m.vy = vy;
        if ((vx * vx) + (vy * vy) > CS) throw Fail("The magnitude of velocity is too high");
    }
}

// --------------------------------------------------------- domain handling --
// GeometryHandler::updateDomainIdentifier (:58-66) -> PhysicalNode.cpp:109-124
static void updateDomainIdentifier(Sim& s) {
    s.reinit_log.clear();
    for (long i = 0; i < (long)s.f.size(); i++) {
        unsigned t = s.dom[i];
        for (int d = 0; d < 5; d++)
            if (s.bn[i][d] >= 0) t = s.bnodes[s.bn[i][d]].domain;
        if (s.dom[i] != t) {
            s.dom[i] = t;
            s.reinit_log.push_back((unsigned)i);
            for (size_t k = 0; k < s.sp.size(); k++) speciesReinit(s, (int)k, i);
        }
    }
}
// GeometryHandler::updateAllDomainIdentifiers (:68-85)
static void updateAllDomainIdentifiers(Sim& s) {
    for (long y = 0; y < s.ny; y++) {
        unsigned cur = s.dom[s.idx(0, y)];
        if (cur != 0) throw Fail("domainID at the left boundary must be 0.");
        for (long x = 1; x < s.nx; x++) {
            const long i = s.idx(x, y);
            if (s.bn[i][W] >= 0) cur = s.bnodes[s.bn[i][W]].domain;
            s.dom[i] = cur;
        }
    }
}
// GeometryHandler::copyCellTypeToPhysicalNodes (:267-301)
static void copyCellType(Sim& s) {
    for (long i = 0; i < (long)s.f.size(); i++) {
        auto it = s.celltype_map.find(s.dom[i]);
        if (it == s.celltype_map.end()) s.celltype_map[s.dom[i]] = 1;
        s.ctype[i] = s.celltype_map[s.dom[i]];
    }
}
// GeometryHandler::moveLattice (:358-383)
static void moveLattice(Sim& s) {
    for (auto& kv : s.markers) {     // Geometry::moveGeometryNodes (Geometry.cpp:454-469)
        Marker& m = *kv.second;
        m.ox = m.x; m.oy = m.y;
        setPos(s, m, m.x + m.vx, m.y + m.vy);
    }
    perturbConnections(s);
    regenerateBoundaryNodes(s);
    connectAll(s);
    updateDomainIdentifier(s);
    updateAllDomainIdentifiers(s);
    copyCellType(s);
}
// GeometryHandler::cureLattice (:149-170)
static void cureLattice(Sim& s) {
    perturbConnections(s);
    connectAll(s);
    regenerateBoundaryNodes(s);
    updateAllDomainIdentifiers(s);
}

// ------------------------------------------------------ topology edits -----
static Marker* addNode(Sim& s, unsigned id, double x, double y) {   // GeometryNode ctor
    auto m = std::unique_ptr<Marker>(new Marker());
    m->id = id; m->x = x; m->y = y; m->ox = x - 5.0; m->oy = y - 5.0;
    m->stencil.fill(0);
    perturbatePosition(s, *m);
    Marker* p = m.get();
    s.markers[id] = std::move(m);
    return p;
}
static void addConnection(Sim& s, Marker* a, Marker* b, unsigned dom, const std::vector<int>& bc) {
    auto c = std::make_shared<Conn>();   // Geometry::addConnection (Geometry.cpp:433-452)
    c->a = a; c->b = b; c->domain = dom; c->bc_species = bc;
    s.conns.push_back(c);
    a->c1 = c.get();
    b->c0 = c.get();
}
// GeometryHandler::remeshBoundary (:385-431).  Each long connection is erased
// from connections_ (order of the others kept) and two halves are appended, in
// the order the copy is visited; one pass with a survivor list gives the same
// vector without the reference's per-erase scans.
static unsigned remeshBoundary(Sim& s) {
    std::vector<std::shared_ptr<Conn>> copy(s.conns);
    std::vector<std::shared_ptr<Conn>> keep, added;
    keep.reserve(copy.size());
    unsigned counter = 0;
    for (auto& c : copy) {
        const double len = std::sqrt((c->a->x - c->b->x) * (c->a->x - c->b->x) +
                                     (c->a->y - c->b->y) * (c->a->y - c->b->y));
        if (len > MAXLENGTH) {
            const double x1 = c->a->x, y1 = c->a->y, x2 = c->b->x, y2 = c->b->y;
            const unsigned nid = s.markers.rbegin()->first + 1;
            Marker* nm = addNode(s, nid, (x1 + x2) / 2, (y1 + y2) / 2);
            Marker* a = c->a; Marker* b = c->b;
            if (a->c1 == c.get()) a->c1 = nullptr;   // Geometry::eraseConnection
            if (b->c0 == c.get()) b->c0 = nullptr;
            for (int h = 0; h < 2; h++) {
                auto n = std::make_shared<Conn>();   // Geometry::addConnection
                n->a = h == 0 ? a : nm; n->b = h == 0 ? nm : b;
                n->domain = c->domain; n->bc_species = c->bc_species;
                n->a->c1 = n.get(); n->b->c0 = n.get();
                added.push_back(n);
            }
            counter++;
        } else {
            keep.push_back(c);
        }
    }
    if (counter) {
        keep.insert(keep.end(), added.begin(), added.end());
        s.conns.swap(keep);
    }
    connectAll(s);
    return counter;
}
// GeometryHandler::coarsenBoundary (:433-461) with Geometry::removeGeometryNode
// (Geometry.cpp:401-428): erased connections are flagged and compacted at the
// end, appended ones keep their creation order.
static void coarsenBoundary(Sim& s) {
    std::vector<std::shared_ptr<Conn>> all(s.conns);
    std::map<const Conn*, bool> dead;
    bool any = false;
    for (auto it = s.markers.begin(); it != s.markers.end();) {
        Marker& m = *it->second;
        const Marker& P1 = *m.c0->a;
        const Marker& P2 = *m.c1->b;
        const double length = std::sqrt((P2.x - P1.x) * (P2.x - P1.x) + (P2.y - P1.y) * (P2.y - P1.y));
        if (length < MAXLENGTH) {
            Conn* C1 = m.c0; Conn* C2 = m.c1;
            Marker* N1 = C1->a; Marker* N2 = C2->b;
            const std::vector<int> bc = C1->bc_species;
            const unsigned dom = C1->domain;     // getDomainIdOfAdjacentConnections
            for (Conn* c : {C1, C2}) {
                if (c->a->c1 == c) c->a->c1 = nullptr;
                if (c->b->c0 == c) c->b->c0 = nullptr;
                dead[c] = true;
            }
            it = s.markers.erase(it);
            auto n = std::make_shared<Conn>();
            n->a = N1; n->b = N2; n->domain = dom; n->bc_species = bc;
            N1->c1 = n.get(); N2->c0 = n.get();
            all.push_back(n);
            any = true;
        } else {
            ++it;
        }
    }
    if (any) {
        std::vector<std::shared_ptr<Conn>> out;
        out.reserve(all.size());
        for (auto& c : all)
            if (!dead.count(c.get())) out.push_back(c);
        s.conns.swap(out);
    }
}

// ---------------------------------------------------------- bio / mass ------
// ForceSolver::addForce via the text descriptor (ForceSolver.cpp:107-139): the
// tension solver streams FORCE with the default 6 significant digits.
static double via_text(double v) {
    std::stringstream ss; ss << v; double r; ss >> r; return r;
}
static void addForceStruct(Sim& s, const ForceStruct& fs) { s.forcemap[fs.id1].push_back(fs); }
static void deleteForceType(Sim& s, int type) {       // ForceSolver.cpp:168-204
    for (auto& kv : s.forcemap)
        kv.second.erase(std::remove_if(kv.second.begin(), kv.second.end(),
                                       [type](const ForceStruct& f) { return f.type == type; }),
                        kv.second.end());
}
// GeometryHandler::computeAreas (:101-134): ring from the first connection of
// each domain, appending `second` points.  Area by the shoelace sum in ring
// order (boost::geometry::area is absent here; parity against it is unpinned).
static std::map<unsigned, double> computeAreas(const Sim& s) {
    std::map<unsigned, std::vector<const Conn*>> def;
    for (auto& c : s.conns) def[c->domain].push_back(c.get());
    std::map<unsigned, double> out;
    for (auto& kv : def) {
        std::vector<double> xs, ys;
        const Conn* start = kv.second[0];
        const Conn* t = start;
        do { xs.push_back(t->b->x); ys.push_back(t->b->y); t = t->b->c1; } while (t != start);
        double sum = 0.0;
        const size_t n = xs.size();
        for (size_t k = 0; k < n; k++) {
            const size_t k1 = (k + 1 == n) ? 0 : k + 1;
            sum += xs[k] * ys[k1] - xs[k1] * ys[k];
        }
        out[kv.first] = sum * 0.5;
    }
    return out;
}
// ------------------------------------------------------- cell division ------
// BioSolverCellDivision / tutorial_0{1,2}_BioSolverCellDivision (…CellDivision.cpp:42-256)
using CellDef = std::map<unsigned, std::vector<std::shared_ptr<Conn>>>;
static CellDef cellDefinition(const Sim& s) {        // updateCellDefinition (:250-256)
    CellDef d;
    for (auto& c : s.conns) d[c->domain].push_back(c);
    return d;
}
static bool belowAxis(const Conn& c, double invslope, double intercept) {   // (:184-190)
    const bool C1 = c.a->y > invslope * c.a->x + intercept;
    const bool C2 = c.b->y < invslope * c.b->x + intercept;
    return (C1 && C2) || (!C1 && !C2);
}
// getTwoConnectionsLongestAxis (:151-198): the farthest pair of first nodes
// (boost::geometry::distance, pythagoras: the sum of two squares is
// commutative, so its term order cannot matter), strict > keeps the first;
// the cut is the perpendicular bisector of that pair.
static std::vector<std::shared_ptr<Conn>> longestAxis(const std::vector<std::shared_ptr<Conn>>& cell) {
    double dist = 0.0;
    Marker *N1 = nullptr, *N2 = nullptr;
    for (auto& c1 : cell)
        for (auto& c2 : cell) {
            const double dx = c1->a->x - c2->a->x, dy = c1->a->y - c2->a->y;
            const double t = std::sqrt(dx * dx + dy * dy);
            if (t > dist) { dist = t; N1 = c1->a; N2 = c2->a; }
        }
    if (!N1) throw Fail("could not find cell division plane");
    const double mx = 0.5 * (N1->x + N2->x), my = 0.5 * (N1->y + N2->y);
    const double invslope = -(N2->x - N1->x) / (N2->y - N1->y);
    const double intercept = my - invslope * mx;
    if (!std::isfinite(invslope) || !std::isfinite(intercept)) throw Fail("could not find cell division plane");
    std::vector<std::shared_ptr<Conn>> out;
    for (auto& c : cell)
        if (belowAxis(*c, invslope, intercept)) out.push_back(c);
    if (out.size() != 2)
        throw Fail("More or less than 2 connections found in BioSolver::getTwoConnectionsLongestAxis().");
    return out;
}
// getTwoConnectionsRandomDirection (:200-248): a line through the mean of the
// first nodes at angle U(0, 2 pi) from the solver's mt19937(1), retried up to
// 10 times until it cuts exactly two connections.
static std::vector<std::shared_ptr<Conn>> randomDirection(const std::vector<std::shared_ptr<Conn>>& cell,
                                                          std::mt19937& gen) {
    std::uniform_real_distribution<> dis(0.0, 2.0 * M_PI);
    double cx = 0.0, cy = 0.0;
    for (auto& c : cell) { cx += c->a->x; cy += c->a->y; }
    cx /= (double)cell.size();
    cy /= (double)cell.size();
    std::vector<std::shared_ptr<Conn>> out;
    unsigned counter = 0;
    while (out.size() != 2) {
        const double invslope = std::tan(dis(gen));
        const double intercept = cy - invslope * cx;
        out.clear();
        for (auto& c : cell)
            if (belowAxis(*c, invslope, intercept)) out.push_back(c);
        counter++;
        if (counter > 10) throw Fail("could not find cell division plane");
    }
    return out;
}
// divideCell (:94-149): seal the two cut connections, close both halves with
// new connections (appended), erase the cut ones, and give the ring through
// the second cut's end node the id max(domain) + 1.
static void divideCell(Sim& s, Sim::Bio& b, unsigned dom, CellDef& def) {
    const unsigned newID = def.rbegin()->first + 1;
    auto aff = (b.kind == 7) ? longestAxis(def[dom]) : randomDirection(def[dom], b.gen);
    aff[0]->a->c1 = nullptr; aff[0]->b->c0 = nullptr;
    aff[1]->a->c1 = nullptr; aff[1]->b->c0 = nullptr;
    if (aff[0]->bc_species != aff[1]->bc_species)
        throw Fail("The method cannot handle cell division cases with heterogeneous boundary conditions.");
    const std::vector<int> bc = aff[0]->bc_species;
    addConnection(s, aff[0]->a, aff[1]->b, dom, bc);
    addConnection(s, aff[1]->a, aff[0]->b, dom, bc);
    for (auto& c : aff) {                  // Geometry::eraseConnection (Geometry.cpp:471-491)
        if (c->a->c1 == c.get()) c->a->c1 = nullptr;
        if (c->b->c0 == c.get()) c->b->c0 = nullptr;
        s.conns.erase(std::remove(s.conns.begin(), s.conns.end(), c), s.conns.end());
    }
    Conn* start = aff[1]->b->c1;
    Conn* c1 = start;
    Conn* c2 = nullptr;
    start->domain = newID;
    while (c2 != start) {
        c2 = c1->b->c1;
        c2->domain = newID;
        c1 = c2;
    }
}

static void applyBio(Sim& s) {
    for (auto& b : s.bio) {
        if (s.it % b.freq != 0) continue;
        if (b.kind == 0) {
            // tutorial_01_BioSolverMembraneTension.cpp:42-77
            deleteForceType(s, 6);
            const double F = via_text(b.p[0]);
            for (auto& c : s.conns) {
                addForceStruct(s, ForceStruct{6, c->a->id, c->b->id, F, 0});
                addForceStruct(s, ForceStruct{6, c->b->id, c->a->id, F, 0});
            }
        } else if (b.kind == 1) {
            // tutorial_01_BioSolverAreaRegulator.cpp:43-100 (TARGETAREA,KP,MAXSOURCE = p[0..2];
            // p[3] > 0 scales the target by the first area seen instead of TARGETAREA)
            auto areas = computeAreas(s);
            for (auto& kv : areas)
                if (!b.target.count(kv.first)) b.target[kv.first] = (b.p[3] > 0) ? b.p[3] * kv.second : b.p[0];
            std::map<unsigned, double> src;
            for (auto& kv : areas) src[kv.first] = b.p[1] * (b.target[kv.first] - areas[kv.first]);
            for (auto& kv : src) {
                if (kv.second > b.p[2]) kv.second = b.p[2];
                else if (kv.second < -b.p[2]) kv.second = -b.p[2];
            }
            for (long i = 0; i < (long)s.f.size(); i++) {
                const unsigned d = s.dom[i];
                if (d != 0) s.madd[i] = (s.ctype[i] == 1) ? b.p[2] : src[d];   // addMass overwrites
            }
        } else if (b.kind == 2) {
            // tutorial_02_BioSolverCellJunction.cpp:47-99 (RADIUS, K, L0 = p[0..2]), one
            // thread: deleteForceType(7); for each node in id order the closest node
            // within RADIUS (FastNeighborList.hpp:404-467: removed when distance >
            // radius) whose getDomainIdOfAdjacentConnections (GeometryNode.cpp:267-283)
            // differs; K and L0 pass through the text descriptor.  Brute force over
            // all nodes in id order, so equidistant candidates resolve to the smaller id.
            deleteForceType(s, 7);
            const double R = b.p[0], K = via_text(b.p[1]), L0 = via_text(b.p[2]);
            // candidates within R lie in the 3x3 bins of side R around a node;
            // visiting them in id order keeps the all-pairs result (first
            // strictly closer node in id order wins)
            std::vector<const Marker*> nodes;
            std::map<std::pair<long, long>, std::vector<const Marker*>> bins;
            auto key = [R](double x, double y) { return std::make_pair((long)std::floor(x / R), (long)std::floor(y / R)); };
            for (auto& kv : s.markers) {
                nodes.push_back(kv.second.get());
                bins[key(kv.second->x, kv.second->y)].push_back(kv.second.get());
            }
            std::vector<const Marker*> cand;
            for (const Marker* m : nodes) {
                const unsigned avoid = m->c0->domain;
                const Marker* best = nullptr;
                double bd = 0.0;
                cand.clear();
                const auto k0 = key(m->x, m->y);
                for (long by = k0.second - 1; by <= k0.second + 1; by++)
                    for (long bx = k0.first - 1; bx <= k0.first + 1; bx++) {
                        auto it = bins.find(std::make_pair(bx, by));
                        if (it != bins.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
                    }
                std::sort(cand.begin(), cand.end(), [](const Marker* a, const Marker* c) { return a->id < c->id; });
                for (const Marker* q : cand) {
                    const double d = std::sqrt((q->x - m->x) * (q->x - m->x) + (q->y - m->y) * (q->y - m->y));
                    if (d > R) continue;
                    if (q->c0->domain == avoid) continue;
                    if (!best || d < bd) { best = q; bd = d; }
                }
                if (best) {
                    addForceStruct(s, ForceStruct{7, m->id, best->id, K, L0});
                    addForceStruct(s, ForceStruct{7, best->id, m->id, K, L0});
                }
            }
        } else if (b.kind == 3) {
            // BioSolverHookeanMembraneTension.cpp:47-80 (K, l_resting = p[0..1]): deletes
            // type 6, adds type-0 springs along both directions of every connection
            deleteForceType(s, 6);
            const double K = via_text(b.p[0]), L = via_text(b.p[1]);
            for (auto& c : s.conns) {
                addForceStruct(s, ForceStruct{0, c->a->id, c->b->id, K, L});
                addForceStruct(s, ForceStruct{0, c->b->id, c->a->id, K, L});
            }
        } else if (b.kind == 4) {
            // BioSolverGrowth.cpp:44-66: addMass(source = p[0]) on every cell site
            for (long i = 0; i < (long)s.f.size(); i++)
                if (s.dom[i] != 0) s.madd[i] = b.p[0];
        } else if (b.kind == 5) {
            // tutorial_02_BioSolverGrowth.cpp:44-82 (MAXSOURCE, THRESHOLD = p[0..1]):
            // R*R*L > THRESHOLD adds MAXSOURCE, otherwise MAXSOURCE/10
            int iR = -1, iL = -1;
            for (size_t k = 0; k < s.sp.size(); k++) {
                if (s.sp[k].kind == SP_SCHNAKENBERG_R) iR = (int)k;
                if (s.sp[k].kind == SP_SCHNAKENBERG_L) iL = (int)k;
            }
            if (iR < 0 || iL < 0) throw Fail("The CDE Solver with the name tutorial_02_CDESolverD2Q5_R does not exists.");
            for (long i = 0; i < (long)s.f.size(); i++) {
                if (s.dom[i] == 0) continue;
                const double R = cOfSlots(s.g[iR][i]), L = cOfSlots(s.g[iL][i]);
                s.madd[i] = (R * R * L > b.p[1]) ? b.p[0] : b.p[0] / 10.0;
            }
        } else if (b.kind == 6) {
            // tutorial_01_BioSolverDifferentiation.cpp:42-83 (THRESHOLD = p[0], species
            // index p[1]): accumulated concentration per domain
            // (GeometryHandler.cpp:136-147, x from 1), tracker reset once, every
            // domain after the first below THRESHOLD becomes type 2
            const int k = (int)b.p[1];
            std::map<unsigned, double> acc;
            for (long y = 0; y < s.ny; y++)
                for (long x = 1; x < s.nx; x++) acc[s.dom[s.idx(x, y)]] += cOfSlots(s.g[k][s.idx(x, y)]);
            if (!b.inited) {
                s.celltype_map.clear();
                s.celltype_map[0] = 0;
                for (auto& kv : cellDefinition(s)) s.celltype_map[kv.first] = 1;
                b.inited = true;
            }
            if (!acc.empty())
                for (auto it = ++acc.begin(); it != acc.end(); ++it)
                    if (it->second < b.p[0]) s.celltype_map[it->first] = 2;
            copyCellType(s);
        } else if (b.kind == 7 || b.kind == 8) {
            // cell division, longest axis (7) or random direction (8); MAXCELLSIZE = p[0]
            CellDef def = cellDefinition(s);
            if (s.celltype_map.empty()) {
                s.celltype_map[0] = 0;
                for (auto& kv : def) s.celltype_map[kv.first] = 1;
            }
            auto areas = computeAreas(s);
            areas[0] = 0;
            for (auto it = ++areas.begin(); it != areas.end(); ++it) {
                if (it->second > b.p[0]) {
                    divideCell(s, b, it->first, def);
                    def = cellDefinition(s);
                    s.celltype_map[def.rbegin()->first] = s.celltype_map[it->first];
                    cureLattice(s);
                }
            }
            copyCellType(s);
        } else if (b.kind == 9) {
            // BioSolverAreaRegulator.cpp:44-130 (TARGETAREA, TARGETAREA2, KP, MAXSOURCE =
            // p[0..3], KI = 0): cell 1 aims at TARGETAREA2 before iteration 5000;
            // sources clamped to [0, MAXSOURCE]; cell type 1 sites take MAXSOURCE
            const double KI = 0.0;
            if (s.it < 5000) { b.target[1] = b.p[1]; b.integ[1] = 0.0; }
            else b.target[1] = b.p[0];
            auto areas = computeAreas(s);
            for (auto& kv : areas)
                if (b.target.find(kv.first) == b.target.end()) { b.target[kv.first] = b.p[0]; b.integ[kv.first] = 0.0; }
            std::map<unsigned, double> err, src;
            for (auto& kv : areas) err[kv.first] = b.target[kv.first] - areas[kv.first];
            for (auto& kv : areas) b.integ[kv.first] += err[kv.first];
            for (auto& kv : areas) {
                const double u = b.p[2] * err[kv.first] + KI * b.integ[kv.first];
                src[kv.first] = (u > b.p[3]) ? b.p[3] : ((u < 0.0) ? 0.0 : u);
            }
            for (long i = 0; i < (long)s.f.size(); i++) {
                const unsigned d = s.dom[i];
                if (d != 0) s.madd[i] = (s.ctype[i] == 1) ? b.p[3] : src[d];
            }
        }
    }
}
// MassSolverBoxOutlet::calculateMass (MassSolverBoxOutlet.cpp:42-72)
static void boxOutlet(Sim& s) {
    auto resc = [&](long i) {
        const double sf = 1.0 / rhoOfSlots(s.f[i]);
        for (auto& v : s.f[i]) v *= sf;
        for (size_t k = 0; k < s.sp.size(); k++) for (auto& v : s.g[k][i]) v *= sf;
    };
    for (long y = 0; y < s.ny; y++) resc(s.idx(0, y));
    for (long x = 0; x < s.nx; x++) resc(s.idx(x, 0));
}

// --------------------------------------------------------------- the step --
static void collide(Sim& s) {       // SimulationRunner.cpp:244-261
    const long n = (long)s.f.size();
#pragma omp parallel for schedule(static) num_threads(s.threads) if (s.threads > 1)
    for (long i = 0; i < n; i++) {
        fluidCollide(s, i);
        double cj[16];
        for (size_t j = 0; j < s.sp.size() && j < 16; j++) cj[j] = cOfSlots(s.g[j][i]);
        for (size_t k = 0; k < s.sp.size(); k++) speciesCollide(s, (int)k, i, cj);
    }
}
// SimulationRunner::runSimulation loop body (SimulationRunner.cpp:72-100),
// with the wall time of each stage accumulated in Sim::stage_s
static void step(Sim& s) {
    auto t = std::chrono::steady_clock::now();
    auto lap = [&](int k) {
        const auto now = std::chrono::steady_clock::now();
        s.stage_s[k] += std::chrono::duration<double>(now - t).count();
        t = now;
    };
    calculateForce(s);
    distributeForce(s);
    lap(0);
    preAdvect(s);
    advect(s);
    postAdvect(s);
    lap(1);
    collide(s);
    lap(2);
    collectVelocity(s);
    moveLattice(s);
    lap(3);
    while (remeshBoundary(s)) {}
    if (s.it % 10 == 0) {
        s.forcemap.clear();
        coarsenBoundary(s);
    }
    lap(4);
    if (s.box_outlet) boxOutlet(s);
    applyBio(s);
    lap(5);
    s.it++;
}

}  // namespace orc

// =============================================================== C ABI ======
using namespace orc;

extern "C" {

void* orc_create(int nx, int ny, double tauf, int force_model) {
    Sim* s = new Sim();
    s->nx = nx; s->ny = ny; s->tauf = tauf; s->force_model = force_model;
    const size_t n = (size_t)nx * ny;
    s->f.assign(n, std::array<double, 9>{});
    s->ux.assign(n, 0); s->uy.assign(n, 0); s->frx.assign(n, 0); s->fry.assign(n, 0);
    s->madd.assign(n, 0); s->dom.assign(n, 0); s->ctype.assign(n, 0);
    s->bn.assign(n, std::array<long, 5>{-1, -1, -1, -1, -1});
    s->celltype_map[0] = 0;   // Geometry::loadGeometryTXT (Geometry.cpp:104)
    return s;
}
void orc_destroy(void* p) { delete static_cast<Sim*>(p); }
const char* orc_last_error(void* p) { return static_cast<Sim*>(p)->last_error.c_str(); }
void orc_set_threads(void* p, int t) { static_cast<Sim*>(p)->threads = t < 1 ? 1 : t; }
void orc_set_iteration(void* p, unsigned it) { static_cast<Sim*>(p)->it = it; }
unsigned orc_iteration(void* p) { return static_cast<Sim*>(p)->it; }

int orc_add_species(void* p, int kind, double tau, const double* prm, int nprm) {
    Sim* s = static_cast<Sim*>(p);
    Species sp; sp.kind = kind; sp.tau = tau;
    if (kind == SP_SIGNAL && nprm >= 5) {
        sp.production = prm[0]; sp.decay = prm[1]; sp.switchoff = (unsigned)prm[2];
        sp.init_inside = prm[3]; sp.prod_domain = (int)prm[4];
    }
    if (kind == SP_LINEAR) {
        // prm = [source_inside, init_inside, k_0 .. k_{ns-1}]
        if (nprm >= 1) sp.source_inside = prm[0];
        if (nprm >= 2) sp.init_inside = prm[1];
        for (int j = 2; j < nprm; j++) sp.krow.push_back(prm[j]);
    }
    if ((kind == SP_SCHNAKENBERG_R || kind == SP_SCHNAKENBERG_L) && nprm >= 5) {
        // prm = [gamma, a or b, deltaT, partner index, init_inside]
        sp.gamma = prm[0]; sp.ab = prm[1]; sp.deltaT = prm[2]; sp.partner = (int)prm[3];
        sp.init_inside = prm[4];
    }
    s->sp.push_back(sp);
    s->g.emplace_back(s->f.size(), std::array<double, 5>{});
    return (int)s->sp.size() - 1;
}
void orc_add_node(void* p, unsigned id, double x, double y) { addNode(*static_cast<Sim*>(p), id, x, y); }
int orc_add_connection(void* p, unsigned id1, unsigned id2, unsigned dom, const int* bc, int nbc) {
    Sim* s = static_cast<Sim*>(p);
    auto a = s->markers.find(id1), b = s->markers.find(id2);
    if (a == s->markers.end() || b == s->markers.end()) return -1;
    std::vector<int> v(bc, bc + nbc);
    addConnection(*s, a->second.get(), b->second.get(), dom, v);
    s->celltype_map[dom] = 1;   // Geometry.cpp:140
    return 0;
}
int orc_cure_lattice(void* p) {
    Sim* s = static_cast<Sim*>(p);
    try { cureLattice(*s); } catch (const std::exception& e) { s->last_error = e.what(); return -1; }
    return 0;
}
void orc_init_solvers(void* p, const double* u0) {   // SimulationRunner.cpp:290-320
    Sim* s = static_cast<Sim*>(p);
    for (long i = 0; i < (long)s->f.size(); i++) {
        if (u0) fluidInitVelocity(*s, i, u0[2 * i], u0[2 * i + 1]);
        else fluidInit(*s, i);
        for (size_t k = 0; k < s->sp.size(); k++) speciesInit(*s, (int)k, i);
    }
}
void orc_add_force(void* p, int type, unsigned id1, unsigned id2, double a, double b, double c, double d) {
    ForceStruct fs{type, id1, id2, a, b, c, d};
    addForceStruct(*static_cast<Sim*>(p), fs);
}
void orc_set_celltype(void* p, unsigned dom, unsigned type) { static_cast<Sim*>(p)->celltype_map[dom] = type; }
void orc_enable_box_outlet(void* p, int on) { static_cast<Sim*>(p)->box_outlet = on != 0; }
void orc_add_bio(void* p, int kind, const double* prm, unsigned freq) {
    Sim::Bio b; b.kind = kind; b.freq = freq;
    for (int k = 0; k < 4; k++) b.p[k] = prm ? prm[k] : 0.0;
    static_cast<Sim*>(p)->bio.push_back(b);
}
// the BioSolvers of the current iteration alone (SimulationRunner.cpp:95-97)
int orc_apply_bio(void* p) {
    Sim* s = static_cast<Sim*>(p);
    try {
        applyBio(*s);
    } catch (const std::exception& e) { s->last_error = e.what(); return -1; }
    return 0;
}
// GeometryHandler::copyCellTypeToPhysicalNodes with the current tracker map
void orc_copy_cell_types(void* p) { copyCellType(*static_cast<Sim*>(p)); }
// FluidSolver::rhoToAdd_ of every site (the mass the next collide adds)
void orc_get_mass(void* p, double* m) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->madd.size(); i++) m[i] = s->madd[i];
}
// ForceSolver::forcemap_ flattened in map order (node id, then row order)
int orc_num_forces(void* p) {
    long n = 0;
    for (auto& kv : static_cast<Sim*>(p)->forcemap) n += (long)kv.second.size();
    return (int)n;
}
void orc_get_forces(void* p, int32_t* type, uint32_t* id1, uint32_t* id2, double* prm) {
    long j = 0;
    for (auto& kv : static_cast<Sim*>(p)->forcemap)
        for (auto& f : kv.second) {
            type[j] = f.type; id1[j] = f.id1; id2[j] = f.id2;
            prm[4 * j] = f.a; prm[4 * j + 1] = f.b; prm[4 * j + 2] = f.c; prm[4 * j + 3] = f.d;
            j++;
        }
}
int orc_run(void* p, int steps) {
    Sim* s = static_cast<Sim*>(p);
    try {
        for (int t = 0; t < steps; t++) step(*s);
    } catch (const std::exception& e) { s->last_error = e.what(); return -1; }
    return 0;
}

// --- lattice-only building blocks (used by closed-form tests) -------------
void orc_set_force_field(void* p, const double* fx, const double* fy) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++) { s->frx[i] = fx[i]; s->fry[i] = fy[i]; }
}
void orc_set_mass(void* p, const double* m) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++) s->madd[i] = m[i];
}
void orc_set_domains(void* p, const unsigned* d) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++) s->dom[i] = d[i];
}
// One advect+collide pass with the current force/mass fields (no IB, no geometry).
void orc_lattice_step(void* p, int steps) {
    Sim* s = static_cast<Sim*>(p);
    for (int t = 0; t < steps; t++) {
        preAdvect(*s); advect(*s); postAdvect(*s); collide(*s);
        if (s->box_outlet) boxOutlet(*s);
        s->it++;
    }
}
// Natural-order post-collision populations: out[q*n + i] (slot d of the stored
// state holds direction INV[d]).
void orc_get_f(void* p, double* out) {
    Sim* s = static_cast<Sim*>(p);
    const size_t n = s->f.size();
    for (size_t i = 0; i < n; i++) for (int q = 0; q < 9; q++) out[INV[q] * n + i] = s->f[i][q];
}
void orc_set_f(void* p, const double* in) {
    Sim* s = static_cast<Sim*>(p);
    const size_t n = s->f.size();
    for (size_t i = 0; i < n; i++) for (int q = 0; q < 9; q++) s->f[i][q] = in[INV[q] * n + i];
}
void orc_get_g(void* p, int k, double* out) {
    Sim* s = static_cast<Sim*>(p);
    const size_t n = s->f.size();
    for (size_t i = 0; i < n; i++) for (int q = 0; q < 5; q++) out[INV[q] * n + i] = s->g[k][i][q];
}
void orc_set_g(void* p, int k, const double* in) {
    Sim* s = static_cast<Sim*>(p);
    const size_t n = s->f.size();
    for (size_t i = 0; i < n; i++) for (int q = 0; q < 5; q++) s->g[k][i][q] = in[INV[q] * n + i];
}
// SimulationRunner::writeSimulation's solver file (SimulationRunner.cpp:123-143)
// from the stored slots: FluidSolver::writeSolver (FluidSolver.cpp:200-206) and
// CDESolverD2Q5::writeSolver (CDESolverD2Q5.cpp:52-58) print "x\ty" and the
// distributions_ array in index order with the stream's default formatting,
// row by row over the grid[y][x] of physical nodes.  species_names: one name
// per line, in the order the species were added.
int orc_write_solver_dump(void* p, const char* file, const char* species_names) {
    Sim* s = static_cast<Sim*>(p);
    std::ofstream out(file);
    if (!out.is_open()) return -1;
    std::vector<std::string> names;
    {
        std::stringstream ss(species_names ? species_names : "");
        std::string nm;
        while (std::getline(ss, nm)) if (!nm.empty()) names.push_back(nm);
    }
    out << "#Solvers (xPos\tyPos\tDistributions...)\n";
    out << "#BEGIN FluidSolver\n";
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++) {
            const auto& d = s->f[(size_t)y * s->nx + x];
            out << x << '\t' << y;
            for (double v : d) out << '\t' << v;
            out << '\n';
        }
    out << "#END FluidSolver\n";
    for (size_t k = 0; k < s->g.size() && k < names.size(); k++) {
        out << "#BEGIN " << names[k] << "\n";
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                const auto& d = s->g[k][(size_t)y * s->nx + x];
                out << x << '\t' << y;
                for (double v : d) out << '\t' << v;
                out << '\n';
            }
        out << "#END " << names[k] << "\n";
    }
    return out.good() ? 0 : -1;
}

void orc_get_velocity(void* p, double* ux, double* uy) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++) { ux[i] = s->ux[i]; uy[i] = s->uy[i]; }
}
void orc_get_force_field(void* p, double* fx, double* fy) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++) { fx[i] = s->frx[i]; fy[i] = s->fry[i]; }
}
void orc_get_domains(void* p, unsigned* d, unsigned* ct) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++) { d[i] = s->dom[i]; if (ct) ct[i] = s->ctype[i]; }
}
// Boundary-neighbour table: per site the domain of the boundary node in slot
// E,N,W,S (or -1).
void orc_get_bn_domains(void* p, int* out) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++)
        for (int d = 1; d < 5; d++) out[4 * i + d - 1] = s->bn[i][d] < 0 ? -1 : (int)s->bnodes[s->bn[i][d]].domain;
}
int orc_num_markers(void* p) { return (int)static_cast<Sim*>(p)->markers.size(); }
int orc_num_connections(void* p) { return (int)static_cast<Sim*>(p)->conns.size(); }
// markers in id order: id, x, y, vx, vy, fx, fy
void orc_get_markers(void* p, unsigned* ids, double* xy, double* v, double* fxy) {
    Sim* s = static_cast<Sim*>(p);
    size_t k = 0;
    for (auto& kv : s->markers) {
        const Marker& m = *kv.second;
        if (ids) ids[k] = m.id;
        if (xy) { xy[2 * k] = m.x; xy[2 * k + 1] = m.y; }
        if (v) { v[2 * k] = m.vx; v[2 * k + 1] = m.vy; }
        if (fxy) { fxy[2 * k] = m.fx; fxy[2 * k + 1] = m.fy; }
        k++;
    }
}
void orc_set_marker_forces(void* p, const double* fxy) {
    Sim* s = static_cast<Sim*>(p);
    size_t k = 0;
    for (auto& kv : s->markers) { kv.second->fx = fxy[2 * k]; kv.second->fy = fxy[2 * k + 1]; k++; }
}
void orc_get_connections(void* p, unsigned* id1, unsigned* id2, unsigned* dom) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t k = 0; k < s->conns.size(); k++) {
        id1[k] = s->conns[k]->a->id; id2[k] = s->conns[k]->b->id; dom[k] = s->conns[k]->domain;
    }
}
int orc_num_areas(void* p) { return (int)computeAreas(*static_cast<Sim*>(p)).size(); }
void orc_get_areas(void* p, unsigned* dom, double* area) {
    auto a = computeAreas(*static_cast<Sim*>(p));
    size_t k = 0;
    for (auto& kv : a) { dom[k] = kv.first; area[k] = kv.second; k++; }
}
// IB building blocks on the current stencils.
void orc_spread(void* p) { distributeForce(*static_cast<Sim*>(p)); }
void orc_calculate_force(void* p) { calculateForce(*static_cast<Sim*>(p)); }
int orc_interpolate(void* p) {
    Sim* s = static_cast<Sim*>(p);
    try { collectVelocity(*s); } catch (const std::exception& e) { s->last_error = e.what(); return -1; }
    return 0;
}
void orc_set_velocity(void* p, const double* ux, const double* uy) {
    Sim* s = static_cast<Sim*>(p);
    for (size_t i = 0; i < s->f.size(); i++) { s->ux[i] = ux[i]; s->uy[i] = uy[i]; }
}
double orc_fastcos(double x) { return fastCos(x); }
double orc_delta(double dx, double dy) { return delta(dx, dy); }
long orc_perturb_counts(void* p, int which) {
    Sim* s = static_cast<Sim*>(p);
    return which == 0 ? s->n_perturb_node : s->n_perturb_conn;
}
// mt19937(1) + uniform_real_distribution(-1e-6,1e-6) stream, for the KAT test.
void orc_rng_draws(int n, double* out) {
    std::mt19937 g(1);
    std::uniform_real_distribution<> d(-PERTURBATION, PERTURBATION);
    for (int k = 0; k < n; k++) out[k] = d(g);
}
unsigned orc_mt19937_default_10000(void) {
    std::mt19937 g;
    unsigned v = 0;
    for (int k = 0; k < 10000; k++) v = (unsigned)g();
    return v;
}
int orc_reinit_count(void* p) { return (int)static_cast<Sim*>(p)->reinit_log.size(); }
void orc_stage_times(void* p, double* out) {
    for (int k = 0; k < 6; k++) out[k] = static_cast<Sim*>(p)->stage_s[k];
}
// Half-open point test of UtilLib::geometry::Rectangle::pointWithinBounds.
int orc_rect_within(double n, double s_, double w, double e, double x, double y) {
    return (x >= w) && (x < e) && (y <= n) && (y > s_);
}

// --- single stages, for the hand-derived tests ------------------------------
// Alive boundary nodes in creation order: site, direction to the physical
// neighbour, domain, generating connection index, and x, y, vx, vy.
int orc_num_bnodes(void* p) {
    int n = 0;
    for (auto& b : static_cast<Sim*>(p)->bnodes) n += b.alive ? 1 : 0;
    return n;
}
void orc_get_bnodes(void* p, int* site, int* dir, unsigned* dom, int* conn, double* xyv) {
    Sim* s = static_cast<Sim*>(p);
    size_t k = 0;
    for (auto& b : s->bnodes) {
        if (!b.alive) continue;
        site[k] = (int)b.site; dir[k] = b.dir; dom[k] = b.domain; conn[k] = b.conn;
        xyv[4 * k] = b.x; xyv[4 * k + 1] = b.y; xyv[4 * k + 2] = b.vx; xyv[4 * k + 3] = b.vy;
        k++;
    }
}
// LagrangianPoint::velocity_ of every marker (id order), as collectVelocity sets it.
void orc_set_marker_velocities(void* p, const double* v) {
    Sim* s = static_cast<Sim*>(p);
    size_t k = 0;
    for (auto& kv : s->markers) { kv.second->vx = v[2 * k]; kv.second->vy = v[2 * k + 1]; k++; }
}
// preAdvect + advect + postAdvect only: afterwards slot q of every site holds
// the population arriving in direction q (read with orc_get_g_slots).
void orc_stream_bc(void* p) {
    Sim* s = static_cast<Sim*>(p);
    preAdvect(*s); advect(*s); postAdvect(*s);
}
// Stored slots as they are: out[q*n + i] = slot q of site i.
void orc_get_g_slots(void* p, int k, double* out) {
    Sim* s = static_cast<Sim*>(p);
    const size_t n = s->f.size();
    for (size_t i = 0; i < n; i++) for (int q = 0; q < 5; q++) out[q * n + i] = s->g[k][i][q];
}
// A boundary node of domain `dom` in slot `slot` (E,N,W,S) of a site, without
// a connection (for CDE re-initialisation cases).
void orc_mark_boundary(void* p, long site, int slot, unsigned dom) {
    Sim* s = static_cast<Sim*>(p);
    BNode b{(double)(site % s->nx), (double)(site / s->nx), dom};
    b.site = site; b.dir = INV[slot];
    s->bnodes.push_back(b);
    s->bn[site][slot] = (long)s->bnodes.size() - 1;
}
void orc_reinit_site(void* p, int k, long site) { speciesReinit(*static_cast<Sim*>(p), k, site); }
int orc_remesh(void* p) { return (int)remeshBoundary(*static_cast<Sim*>(p)); }
void orc_coarsen(void* p) { coarsenBoundary(*static_cast<Sim*>(p)); }

}  // extern "C"
