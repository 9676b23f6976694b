/*
 * lbib.h -- C ABI of the MI355X (gfx950) per-timestep lattice path of an
 * LBIBCell-compatible tissue simulator.
 *
 * The library owns all device memory of a context (one HIP stream per
 * context).  Every entry point takes and returns plain host pointers and
 * sizes, never throws, and reports failure as a negative status with a message
 * in lbib_last_error().  Population and species buffers exchanged with the
 * host are in natural direction order, direction-major: f[q * n + y * nx + x],
 * q = T,E,N,W,S,NE,NW,SW,SE (LbmLib/include/Direction.hpp:34-44), holding
 * post-collision values (the reference keeps the same values in the "locally
 * swapped" slot order, FluidSolver.cpp:259-264).
 *
 * Citations are relative to the reference repository root; "LbmLib/" means
 * libs/LbmLib/.
 */
#ifndef LBIB_H
#define LBIB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBIB_MAX_SPECIES 4

typedef struct lbib_ctx lbib_ctx;

/* Storage precision of populations. */
enum lbib_precision {
    LBIB_FP64 = 0,        /* FP64 storage and arithmetic (parity path) */
    LBIB_FP32 = 1,        /* FP32 storage of f_i - w_i, FP32 arithmetic */
    LBIB_FP16_DEV = 2,    /* FP16 storage of f_i - w_i*rho_ref, FP32 arithmetic */
    LBIB_BF16_DEV = 3,    /* BF16 storage of f_i - w_i*rho_ref, FP32 arithmetic */
    LBIB_FP16_REM = 4     /* FP32 moments (rho - 1, j; species C) plus FP16 remainders
                             f_i - feq_i(moments), FP32 arithmetic */
};

/* Fluid force model: Luo 1993 is what the reference compiles
 * (LbmLib/include/solver/FluidSolver/FluidSolver.hpp:41). */
enum lbib_force_model {
    LBIB_FORCE_LUO = 0,
    LBIB_FORCE_GUO_AS_WRITTEN = 1,  /* GuoZhengShi2002ForceModel.cpp:29-78, incl. its NE y-term */
    LBIB_FORCE_GUO = 2              /* Guo et al. 2002 with (9ux + 6uy + 3) in the NE y-term */
};

/* Reaction specs served by the species registry (CDESolverFactory names). */
enum lbib_species_kind {
    LBIB_SPECIES_PLAIN = 0,   /* CDESolverD2Q5 (CDESolverD2Q5.cpp:37-228) */
    LBIB_SPECIES_SIGNAL = 1,  /* tutorial_01_CDESolverD2Q5_SIGNAL (..._SIGNAL.cpp:39-250) */
    LBIB_SPECIES_LINEAR = 2,  /* R_s = sum_j K[s][j] C_j + source inside cells */
    /* tutorial_02_CDESolverD2Q5_R / _L (Schnakenberg pair, ..._R.cpp:38-42,
     * 98-110; ..._L.cpp:38-42, 94-106): inside cells
     *   R: gamma (a - R + R^2 L),  L: gamma (b - R^2 L),
     * added as deltaT*react/3 (rest) and /6 (moving).  Spec fields:
     * production = gamma, source_inside = a (R) or b (L), decay = deltaT,
     * prod_domain = index of the partner species, init_inside = C inside
     * cells at init.  The partner's C is read as the reference reads it:
     * after its collide when it comes earlier in the species order. */
    LBIB_SPECIES_SCHNAKENBERG_R = 3,
    LBIB_SPECIES_SCHNAKENBERG_L = 4
};

typedef struct lbib_species_spec {
    int32_t kind;             /* lbib_species_kind */
    double tau;
    /* SIGNAL: production rate in domain prod_domain (-1: any cell) while
     * it < switchoff, otherwise decay*C sink; init_inside = C inside cells. */
    double production, decay, init_inside;
    uint32_t switchoff;
    int32_t prod_domain;
    /* LINEAR: source inside cells and coupling row K[s][0..nspecies-1]. */
    double source_inside;
    double krow[LBIB_MAX_SPECIES];
} lbib_species_spec;

typedef struct lbib_config {
    int32_t nx, ny;           /* Parameters SizeX, SizeY */
    double tau_fluid;         /* Parameters tauFluid */
    int32_t force_model;      /* lbib_force_model */
    int32_t precision;        /* lbib_precision */
    int32_t nspecies;         /* 0..LBIB_MAX_SPECIES */
    lbib_species_spec species[LBIB_MAX_SPECIES];
    int32_t box_outlet;       /* MassSolverBoxOutlet registered */
    int32_t max_markers;      /* capacity hints (0 = defaults) */
    int32_t max_connections;
} lbib_config;

/* Status word bits returned by lbib_read_status(). */
#define LBIB_ERR_VELOCITY      0x00000001u  /* LagrangianPoint.cpp:96-102 */
#define LBIB_ERR_EDGE          0x00000002u  /* GeometryHandler.cpp:618-636 */
#define LBIB_ERR_SCANLINE      0x00000004u  /* GeometryHandler.cpp:74-76 */
#define LBIB_ERR_SLOPE         0x00000008u  /* Connection.cpp:315-323 */
#define LBIB_ERR_CAPACITY      0x00000010u  /* a device work list overflowed */
#define LBIB_ERR_TOPOLOGY      0x00000020u  /* a domain is not one closed ring (or ring > 4096 markers) */
#define LBIB_ERR_MASK          0x000000ffu
#define LBIB_HALT_PERTURB_NODE 0x00000100u  /* a moved marker needs GeometryNode.cpp:153-214 */
#define LBIB_HALT_PERTURB_CONN 0x00000200u  /* a connection needs Connection.cpp:332-387 */
#define LBIB_HALT_PERTURB_REMESH 0x00000400u /* a remesh midpoint needs GeometryNode.cpp:153-214 */
#define LBIB_HALT_GROW         0x00000800u  /* marker/connection capacity too small for remeshing */
#define LBIB_HALT_MASK         0x00000f00u
#define LBIB_FLAG_REMESH       0x00001000u  /* a segment is longer than MAXLENGTH */
#define LBIB_FLAG_COARSEN      0x00002000u  /* a marker's neighbours are closer than MAXLENGTH */
#define LBIB_SKIPPED           0x00004000u  /* a later step began while stopped: its kernels did nothing */

typedef struct lbib_status {
    uint32_t bits;
    int32_t first_bad_marker;   /* dense marker index of the first error, or -1 */
    int32_t n_touched_sites;    /* 16x16 site tiles that received IB force this step */
    int32_t n_boundary_nodes;   /* cut links kept after closest-wins */
    int32_t n_reinit;           /* sites whose domain changed (CDE reinit) */
    int32_t n_candidates;       /* cut-link candidates generated */
    int32_t n_markers;          /* GeometryNodes after this step's remesh/coarsen */
    int32_t n_connections;      /* Connections after this step's remesh/coarsen */
    int32_t n_remeshed;         /* connections split by remeshing this step */
    int32_t n_coarsened;        /* markers removed by coarsening this step */
    uint32_t step_iteration;    /* iteration of the last lbib_step_begin that found the status clear */
    int32_t n_reinit_levels;    /* sequential levels of the CDE re-initialisation (dependency depth) */
} lbib_status;

/* ---- context ---------------------------------------------------------- */
/* Fails (-1, lbib_last_error) unless 4 <= nx, ny <= 65535, nx * ny <= 2^30
 * (the cut-link selection indexes (site, slot) pairs in 32 bits),
 * 0 <= nspecies <= LBIB_MAX_SPECIES and tau_fluid > 0.5. */
int lbib_create(const lbib_config* cfg, lbib_ctx** out);
/* Replaces: SimulationRunner ctor (SimulationRunner.cpp:50-64) + the per-site
 * solver objects of GeometryHandler::generatePhysicalGrid (GeometryHandler.cpp:463-474). */
void lbib_destroy(lbib_ctx* ctx);
const char* lbib_last_error(const lbib_ctx* ctx);
int lbib_synchronize(lbib_ctx* ctx);
/* Device bytes held by the context. */
size_t lbib_device_bytes(const lbib_ctx* ctx);

/* ---- lattice state ----------------------------------------------------- */
/* FluidSolver::initSolver / initWithVelocity (FluidSolver.cpp:50-101) and the
 * species' initSolver (CDESolverD2Q5.cpp:37-40, ..._SIGNAL.cpp:46-57) on every
 * site; u0 (2*n, interleaved ux,uy) may be NULL.  Species init reads the
 * current domain identifiers. */
int lbib_init_solvers(lbib_ctx* ctx, const double* u0);
/* Whole-field copies (natural order, see top).  FP16/BF16 contexts convert. */
int lbib_set_populations(lbib_ctx* ctx, const double* f);
int lbib_get_populations(lbib_ctx* ctx, double* f);
int lbib_set_species(lbib_ctx* ctx, int s, const double* g);
int lbib_get_species(lbib_ctx* ctx, int s, double* g);
/* FluidSolver::getVelocity for every site (FluidSolver.cpp:242-246). */
int lbib_get_velocity(lbib_ctx* ctx, double* ux, double* uy);
int lbib_get_domains(lbib_ctx* ctx, uint32_t* dom);
int lbib_set_domains(lbib_ctx* ctx, const uint32_t* dom);
/* Prescribed body force field (lattice-only use; the IB spread overwrites it). */
int lbib_set_force_field(lbib_ctx* ctx, const double* fx, const double* fy);
int lbib_get_force_field(lbib_ctx* ctx, double* fx, double* fy);
/* FluidSolver::addMass on every site (FluidSolver.cpp:253-257); consumed and
 * cleared by the next collide. */
int lbib_set_mass_field(lbib_ctx* ctx, const double* m);

/* ---- membrane (GeometryNode) and connections ------------------------------
 * Markers are dense indices 0..nm-1 in increasing GeometryNode id order (the
 * std::map order of Geometry::geometryNodes_).  xy, vxy, oxy are interleaved.
 * lbib_set_markers rejects (status -1, LBIB_ERR_EDGE with first_bad_marker in
 * the status word) any position that is not finite or lies within 2 sites of
 * the lattice edge, or of its rectangle's edge in a batched atlas: the check
 * GeometryHandler::connectGeometryNodesToPhysicalNodes makes for every new
 * node (GeometryHandler.cpp:613-636).  After a change of the marker count,
 * upload ids, connections and forces for the new count. */
int lbib_set_markers(lbib_ctx* ctx, int nm, const double* xy, const double* vxy, const double* oxy);
int lbib_get_markers(lbib_ctx* ctx, double* xy, double* vxy, double* fxy, double* oxy);
/* The markers whose current position meets a condition of
 * GeometryNode::perturbatePosition (GeometryNode.cpp:153-214): the node lies
 * within kEPSILON of an integer coordinate or of a lattice-cell diagonal.
 * *count receives how many there are; the first cap of them (in no set order)
 * get their dense index, GeometryNode id and position written.  Together with
 * lbib_set_marker_positions it replaces the whole-geometry download and upload
 * around GeometryNode::setPos's perturbation when the device halts with
 * LBIB_HALT_PERTURB_NODE: only the flagged nodes draw perturbations. */
int lbib_perturbation_candidates(lbib_ctx* ctx, int cap, int* count, int32_t* index, uint32_t* id, double* xy);
/* The rings (membranes, one per domain id) that hold a connection meeting the
 * condition of Connection::perturbConnection (Connection.cpp:332-387): its two
 * markers within kEPSILON in x or in y.  *nrings receives how many there are;
 * when that is at most cap_rings, domain[0..nrings) lists them in increasing
 * domain id (the order GeometryHandler::perturbConnections visits rings,
 * GeometryHandler.cpp:326-356), *nnodes the total of their markers, and
 * off[0..nrings] the offsets of each ring's markers; when *nnodes is at most
 * cap_nodes, marker[] and xy[] hold each ring's markers (dense index and
 * position) in the walk order of perturbConnections: node j of a ring is the
 * first marker of its j-th connection counted from the ring's first
 * connection in vector order.  Rings are disjoint in markers, so walking
 * these rings alone draws the same perturbations as walking all of them. */
int lbib_perturbation_rings(lbib_ctx* ctx, int cap_rings, int cap_nodes, int* nrings, int* nnodes,
                            uint32_t* domain, int32_t* off, int32_t* marker, double* xy);
/* New positions of n markers given by dense index; no other marker changes.
 * The positions pass the edge check of lbib_set_markers (first_bad_marker is
 * the failing dense index). */
int lbib_set_marker_positions(lbib_ctx* ctx, int n, const int32_t* index, const double* xy);
/* Connections in Geometry::connections_ vector order: first/second marker
 * index, domain id, boundary-condition descriptor index (-1 = none). */
int lbib_set_connections(lbib_ctx* ctx, int nc, const int32_t* a, const int32_t* b,
                         const uint32_t* dom, const int32_t* bc);
/* Boundary-condition descriptors: descriptor d lists the species indices that
 * BoundarySolverNoFluxD2Q5 is connected to (BoundaryNode.cpp:43-46). */
int lbib_set_bc_descriptors(lbib_ctx* ctx, int nd, const int32_t* off, const int32_t* species);
/* Capacity for markers and connections (device topology grows by remeshing).
 * Growth keeps the uploaded topology (positions, ids, links, connections and
 * the force offsets). */
int lbib_reserve(lbib_ctx* ctx, int max_markers, int max_connections);
/* GeometryNode ids per dense marker index (needed once remeshing has added
 * markers on the device) and the current connection list. */
int lbib_get_marker_ids(lbib_ctx* ctx, uint32_t* ids);
int lbib_set_marker_ids(lbib_ctx* ctx, const uint32_t* ids);
int lbib_get_connections(lbib_ctx* ctx, int32_t* a, int32_t* b, uint32_t* dom, int32_t* bc);
/* The force table in the CSR layout of lbib_set_forces (off has nm+1 entries). */
int lbib_get_forces(lbib_ctx* ctx, int32_t* off, int32_t* type, int32_t* partner, double* prm);
/* Cell type per domain id (GeometryHandler::cellTypeTrackerMap_). */
int lbib_set_celltypes(lbib_ctx* ctx, int ndom, const uint32_t* celltype);
/* Force structs of ForceSolver::forcemap_ (ForceSolver.cpp:42-85) in CSR by
 * dense marker index, in map/push order.  prm holds 4 doubles per struct:
 * type 0/7 (k,l0), 5/6 (f), 1 (x,y,k,l0), 2 (fx,fy), 3 (y,k,l0), 4 (x,k,l0). */
int lbib_set_forces(lbib_ctx* ctx, int nm, const int32_t* off, const int32_t* type,
                    const int32_t* partner, const double* prm);

/* ---- the device-resident step ----------------------------------------------
 * Stage A (SimulationRunner.cpp:73-79): calculateForce, distributeForce,
 * pre/postAdvect + advect + collide (+ the MassSolverBoxOutlet rescale), and
 * collectVelocity + the marker move of Geometry::moveGeometryNodes.
 * it = Parameters.currentIteration at this step. */
int lbib_step_lattice(lbib_ctx* ctx, uint32_t it);
/* Stage B (GeometryHandler::moveLattice after the move, :361-382):
 * connection checks, boundary-node generation with closest-wins, domain
 * update + CDE reinit, domain scanline; then remeshBoundary
 * (GeometryHandler.cpp:385-431) and, when it % 10 == 0, deleteAllForces and
 * coarsenBoundary (:433-461) on the device-resident topology. */
int lbib_step_geometry(lbib_ctx* ctx, uint32_t it, int coarsen);
/* Only the remesh/coarsen part of stage B (after a host-handled halt). */
int lbib_step_topology(lbib_ctx* ctx, int coarsen);
/* GeometryHandler::cureLattice (GeometryHandler.cpp:149-170) after the host
 * has perturbed connections: boundary nodes + scanline, no CDE reinit. */
int lbib_cure_lattice(lbib_ctx* ctx);
/* Registers or removes MassSolverBoxOutlet (SimulationRunner::addMassSolver,
 * SimulationRunner.cpp:401-403). */
int lbib_set_box_outlet(lbib_ctx* ctx, int on);
/* Lattice-only step with the current force field (no IB, no geometry). */
int lbib_step_lattice_only(lbib_ctx* ctx, uint32_t it);
/* Synchronises and returns the status of the last stages; clears HALT bits. */
int lbib_read_status(lbib_ctx* ctx, lbib_status* st);
int lbib_clear_status(lbib_ctx* ctx);

/* ---- run-ahead ---------------------------------------------------------------
 * No counterpart in the reference, whose loop (SimulationRunner.cpp:72-100)
 * inspects its geometry after every iteration.  Here a caller may enqueue
 * several iterations and read the status once.  The status is "stopped" when
 * it holds an ERR or HALT bit, LBIB_FLAG_REMESH or LBIB_SKIPPED: then every
 * lattice, IB, force and area-regulator kernel returns without work, and the
 * geometry kernels do so on all of these but LBIB_FLAG_REMESH (the rest of
 * that iteration's topology stage still runs, as it does without run-ahead).
 *
 * lbib_step_begin(ctx, it) opens iteration it.  On the device it records it
 * as lbib_status.step_iteration if the status is clear, and otherwise sets
 * LBIB_SKIPPED, so nothing after the stop changes device state.  On the host
 * it snapshots the launch state (ping-pong buffer, pending mass table,
 * profiling events).  After a stop in iteration t, lbib_rewind(ctx, t + 1)
 * restores the snapshot taken when iteration t + 1 began (and is a no-op if
 * it never began); the caller then resolves iteration t exactly as without
 * run-ahead (host halts, remesh rounds, forces, BioSolvers) and continues
 * with t + 1.  Snapshots of the last 1024 iterations are kept. */
int lbib_step_begin(lbib_ctx* ctx, uint32_t it);
int lbib_rewind(lbib_ctx* ctx, uint32_t it);

/* ---- device BioSolver helpers ------------------------------------------- */
/* Areas (shoelace, FP64, in ring order) per domain id 0..ndom-1; domains
 * without connections read 0.  ndom = value given to lbib_set_celltypes. */
int lbib_compute_areas(lbib_ctx* ctx, double* area);
/* tutorial_01_BioSolverAreaRegulator::applyBioProcess
 * (tutorial_01_BioSolverAreaRegulator.cpp:43-100) as one device op:
 * areas -> clamped proportional source per domain -> addMass on every site of
 * a cell (MAXSOURCE for cell type 1).  target_factor > 0 sets the target area
 * of a new domain to factor * its first area instead of target_area. */
/* GeometryHandler::computeAccumulatedDomainConcentrations
 * (GeometryHandler.cpp:136-147): out[d] = sum of getC() of species s over the
 * sites of domain d (d < ndom), x from 1 as the reference loops.  The sums
 * are accumulated per run of equal domain along rows, so the rounding order
 * differs from the reference's site-by-site sum (relative difference at the
 * 1e-15 level).  Synchronises. */
int lbib_domain_concentrations(lbib_ctx* ctx, int s, int ndom, double* out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory or wait for the device (code under HIP graph capture, serving loops):
 * the same sums as lbib_domain_concentrations, accumulated in a static device
 * scratch array of the library (65536 doubles, one per GPU, held until the
 * process ends) and written to d_out[0 .. ndom), a DEVICE array of the caller,
 * on the context's stream: no allocation, no host copy, no synchronisation
 * (read d_out after lbib_synchronize or a stream-ordered consumer).  ndom is at
 * most 65536 (-1 otherwise).  All calls on one GPU share its scratch array:
 * calls that overlap on a GPU, from any context or thread, race. */
int lbib_domain_concentrations_noalloc(lbib_ctx* ctx, int s, int ndom, double* d_out);
int lbib_area_regulator(lbib_ctx* ctx, double target_area, double kp, double maxsource,
                        double target_factor);
/* tutorial_01_BioSolverMembraneTension (…MembraneTension.cpp:42-77) as one
 * device op: deleteForceType(6) + two type-6 structs of magnitude f per
 * connection.  Only valid while the force table holds no other types. */
int lbib_forces_tension(lbib_ctx* ctx, double f);
/* BioSolverHookeanMembraneTension (BioSolverHookeanMembraneTension.cpp:47-80)
 * as one device op: two type-0 springs (k, l0) per connection, one per
 * direction.  The reference deletes type 6 and appends; this op rewrites the
 * table, so it is valid only while the table is empty (after
 * ForceSolver::deleteAllForces). */
int lbib_forces_springs(lbib_ctx* ctx, double k, double l0);
/* tutorial_02_BioSolverGrowth (tutorial_02_BioSolverGrowth.cpp:44-82) as one
 * device op: on every cell site addMass(R*R*L > threshold ? maxsource :
 * maxsource / 10), R and L the concentrations (getC) of species species_r and
 * species_l; other sites keep this iteration's earlier addMass. */
int lbib_mass_growth(lbib_ctx* ctx, int species_r, int species_l, double threshold, double maxsource);
/* tutorial_02_BioSolverCellJunction / BioSolverCellJunction
 * (tutorial_02_BioSolverCellJunction.cpp:47-99) as one device op:
 * deleteForceType(7), then for each marker j in id order the closest marker c
 * within `radius` whose adjacent connections belong to another domain
 * (Geometry.cpp:351-373, FastNeighborList.hpp:404-467) gives the springs
 * (7, j, c, k, l0) and (7, c, j, k, l0), appended to the rows of j and c.
 * Other struct types keep their order.  k and l0 are used as given (the
 * reference streams them through a text descriptor: round them first).
 * Equidistant candidates resolve to the smaller id.  At most 16 markers may
 * choose one marker; more, or a table beyond its capacity, sets
 * LBIB_ERR_CAPACITY.  Stream-ordered. */
int lbib_forces_junctions(lbib_ctx* ctx, double radius, double k, double l0);
/* ForceSolver::deleteAllForces (ForceSolver.cpp:158-166). */
int lbib_forces_clear(lbib_ctx* ctx);
/* FluidSolver::addMass for every site of each domain: m_site = table[dom]. */
int lbib_add_mass_by_domain(lbib_ctx* ctx, int ndom, const double* table);
/* FluidSolver::addMass (FluidSolver.cpp:253-257) at count listed sites
 * (y * nx + x): each call overwrites the site's mass for the next collide;
 * sites not listed keep what earlier calls of this iteration gave them (a
 * per-domain table from lbib_add_mass_by_domain or a device BioSolver op, or
 * zero).  A site listed twice takes its last value.  This is what a host-side
 * BioSolver's getPhysicalNodes()[y][x]->getFluidSolver().addMass(m) reaches. */
int lbib_add_mass_sites(lbib_ctx* ctx, int64_t count, const int64_t* site, const double* m);

/* ---- batched ensemble -------------------------------------------------------
 * No single counterpart in the reference: a parameter sweep there is one
 * SimulationRunner (SimulationRunner.hpp:51-119) per domain, each a periodic
 * nx x ny lattice (GeometryHandler.cpp:520-589).  Here many such domains share
 * one context: the lattice becomes an atlas in which domain r owns the
 * rectangle [x0, x0+w) x [y0, y0+h) and is periodic inside it, with its own
 * fluid and species relaxation times.  Every per-step launch then serves all
 * domains at once.  x0 and y0 must be multiples of 16 and no 16x16 tile may
 * be shared by two rectangles; sites outside every rectangle are never
 * updated.  Cells of domain r must stay at least 3 sites inside its
 * rectangle (the delta stencil and the cut links never wrap), and their
 * domain identifiers must differ from those of every other domain; the
 * box outlet then acts on each rectangle's x = x0 column and y = y0 row.
 * nrect = 0 returns to one periodic nx x ny domain.  Call after lbib_create
 * and before lbib_init_solvers. */
typedef struct lbib_rect {
    int32_t x0, y0, w, h;
    double tau_fluid;
    double tau_species[LBIB_MAX_SPECIES];
} lbib_rect;
int lbib_set_batch_layout(lbib_ctx* ctx, int nrect, const lbib_rect* rects);

/* Lattice launch shape.  variant 0 = one site per lane in a row tile (the
 * FP64 kernel, and the mixed kernel with sites_per_lane 1, 2 or 4 at
 * workgroup widths 64, 128 or 256); variant 1 = two adjacent sites per lane
 * with 32-bit accesses (16-bit storage, even widths only); variant 2 =
 * sites_per_lane consecutive sites per lane, one wide access per plane (2
 * or 4 in FP32, 4 or 8 in 16-bit storage, widths a multiple of it; 1 for
 * the moments-and-remainders store, whose variant 2 shares loaded moments
 * between lanes).  The default is {0, 256, 1} on one lattice and the paired
 * or two-site shape on a batched atlas.  lbib_autotune times every shape
 * that applies on the current state (scratch velocity output, mass and
 * status untouched: the simulation is not advanced; the force field is read
 * at every site), keeps the fastest and reports all timings;
 * lbib_set_lattice_shape fixes one. */
typedef struct lbib_lattice_shape {
    int32_t variant, workgroup, sites_per_lane;
    double ms;                  /* average launch time measured by the tuner */
} lbib_lattice_shape;
int lbib_autotune(lbib_ctx* ctx, int reps, lbib_lattice_shape* tried, int max_tried, int* n_tried,
                  lbib_lattice_shape* chosen);
int lbib_set_lattice_shape(lbib_ctx* ctx, const lbib_lattice_shape* shape);

/* Step graphs.  No counterpart in the reference (its loop is a host loop,
 * SimulationRunner.cpp:69-110).  Between lbib_capture_begin and
 * lbib_capture_end every per-step call (lbib_step_begin, lbib_step_lattice,
 * lbib_step_geometry, the force and BioSolver operations) is recorded into a
 * HIP graph instead of being run; recorded kernels take the iteration from the
 * device status (one more than the last iteration begun), so the same graph
 * serves every block of steps with the same launch state.  A call that would
 * wait for the device or allocate while recording fails, and lbib_capture_end
 * then reports *ok = 0 and keeps nothing (nothing of the block ran: the caller
 * rewinds with lbib_rewind and runs it directly).  lbib_graph_launch replays
 * the recorded block as iterations it_first, it_first + 1, ... (one launch;
 * run-ahead snapshots are kept as for directly enqueued steps); *stale = 1
 * and nothing is launched when the launch state has changed since recording
 * (a reallocation, another launch shape, a topology upload ...). */
int lbib_capture_begin(lbib_ctx* ctx);
int lbib_capture_end(lbib_ctx* ctx, int* ok);
int lbib_graph_launch(lbib_ctx* ctx, uint32_t it_first, int* stale);
int lbib_graph_reset(lbib_ctx* ctx);

/* ---- adjoint ----------------------------------------------------------------
 * No counterpart in the reference (it has no derivatives).  Reverse-mode
 * gradient of L = sum_x (C_T(x) - C*(x))^2, the final D2Q5 concentration
 * against a target, through the coupled step restricted to its continuous
 * part on one periodic nx x ny lattice: type-6 membrane tension of stiffness
 * kappa on closed marker rings (ForceStructs.hpp:455-514), fastCos-delta
 * spreading and interpolation with X += U (GeometryNode.cpp:55-254), D2Q9 BGK
 * with Luo forcing and no mass source (FluidSolver.cpp:103-183,
 * Luo1993ForceModel.cpp:29-70), and one D2Q5 species on the fresh velocity
 * with reaction rates[c] inside cell c and -decay*C outside
 * (tutorial_01_CDESolverD2Q5_SIGNAL.cpp:92-135, one rate per cell).  The cell
 * map is fixed; cut links, re-initialisation and remeshing are discrete and
 * not part of the model.  Gradients: dL/dtau_c, dL/dkappa, dL/drates[c].
 * States are checkpointed every checkpoint_every steps; lbib_adj_gradient
 * recomputes one segment at a time from its checkpoint and runs the adjoint
 * kernels backwards over it.  precision: LBIB_FP32 or LBIB_FP64 arithmetic
 * and storage (gradients are accumulated in FP64 either way). */
typedef struct lbib_adj lbib_adj;
typedef struct lbib_adj_config {
    int32_t nx, ny, precision;
    double tau_fluid, decay;
    int32_t n_markers, n_cells, checkpoint_every;
} lbib_adj_config;
int lbib_adj_create(const lbib_adj_config* cfg, lbib_adj** out);
void lbib_adj_destroy(lbib_adj* adj);
const char* lbib_adj_last_error(const lbib_adj* adj);
/* Initial populations f0 (9*n) and g0 (5*n) in natural direction order,
 * plane by plane; marker positions xy0 (2*n_markers); next[k] = the ring
 * successor of marker k (closed rings); dom (n) = cell of each site, 0
 * outside, at most n_cells; target (n) = C*. */
int lbib_adj_set_problem(lbib_adj* adj, const double* f0, const double* g0, const double* xy0, const int32_t* next,
                         const uint32_t* dom, const double* target);
/* Runs steps forward steps with the given parameters (rates has n_cells + 1
 * entries, [0] unused), storing checkpoints; returns the loss and, when
 * c_final != NULL, the final concentration (n). */
int lbib_adj_forward(lbib_adj* adj, int steps, double tau_c, double kappa, const double* rates, double* loss,
                     double* c_final);
/* Gradient of the last forward's loss (d_rates: n_cells + 1 entries, [0] = 0). */
int lbib_adj_gradient(lbib_adj* adj, double* d_tau_c, double* d_kappa, double* d_rates);
/* Population storage of the adjoint's states (call before
 * lbib_adj_set_problem): the arithmetic precision (default), or with FP32
 * arithmetic LBIB_FP16_DEV / LBIB_BF16_DEV, 16-bit deviations from the rest
 * state (fluid f - w_q, species as they are) like the mixed-precision lattice;
 * the reverse sweep reads the stored states and takes the rounding's
 * derivative as 1.  No reference counterpart: part of the adjoint. */
int lbib_adj_set_storage(lbib_adj* a, int storage);
/* Checkpointing: LBIB_ADJ_CHECKPOINT_EVERY, param = K (states every K steps,
 * a K-state segment store; the default with config.checkpoint_every), or
 * LBIB_ADJ_REVOLVE, param = s snapshot slots (binomial checkpointing,
 * Griewank & Walther, ACM TOMS 26 (2000), Algorithm 799: T steps with every
 * step recomputed at most r times when T <= C(s + r, s)).  Both give the same
 * gradient; they trade device memory for recomputation. */
#define LBIB_ADJ_CHECKPOINT_EVERY 0
#define LBIB_ADJ_REVOLVE 1
int lbib_adj_set_checkpointing(lbib_adj* a, int scheme, int param);
/* Forward steps the last forward and gradient calls ran (the gradient's count
 * includes every recomputed step) and the bytes of the checkpoint store. */
int lbib_adj_stats(const lbib_adj* a, int64_t* forward_steps, int64_t* gradient_steps, int64_t* store_bytes);
/* Device time of the last forward and of the last gradient (recompute + reverse). */
int lbib_adj_timing(lbib_adj* adj, double* forward_ms, double* gradient_ms);
/* Multi-species model (after lbib_adj_set_problem; replaces its species):
 * nspecies (1-4) species with initial populations g0 (nspecies*5*n), targets
 * (nspecies*n; the loss sums over species) and decay[s] outside every cell.
 * The reaction is R_s = sum_j K_sj C_j (all concentrations before the
 * collides, as the linear-matrix species of the multi-species runs), plus
 * rates[c] inside cell c for species 0, plus -decay_s C_s outside cells. */
int lbib_adj_set_species(lbib_adj* adj, int nspecies, const double* g0, const double* target, const double* decay);
/* Junction springs between marker pairs (pairs: 2*npairs marker indices, each
 * pair once): F_a += k_j (|d| - l0) d/|d| with d = X_b - X_a, and the
 * opposite force on b (ForceType7, ForceStructs.hpp:521-588, as the
 * cell-junction BioSolver creates them); the pair list is fixed over the run. */
int lbib_adj_set_junctions(lbib_adj* adj, int npairs, const int32_t* pairs, double l0);
/* A batch of independent problems in one lattice: nrect periodic rectangles
 * (rects: x0, y0, w, h each; origins on the 16-site tile grid, no 16x16 tile
 * shared), each holding its own cells (every marker inside its rectangle).
 * Parameters are shared by the batch and the loss sums over it, so the
 * gradient is the sum of the problems' gradients; sites outside every
 * rectangle are not updated.  Each marker keeps the rectangle it starts in,
 * and its delta stencil wraps inside that rectangle, as the lattice wraps each
 * problem, so a marker drifting to its rectangle's edge never touches another
 * problem.  Call after lbib_adj_set_problem (which drops the layout); nrect = 0
 * returns to one periodic lattice. */
int lbib_adj_set_batch_layout(lbib_adj* adj, int nrect, const int32_t* rects);
/* Forward with every parameter: tau_c[nspecies], kappa, the junction
 * stiffness k_j, K (nspecies*nspecies row-major, NULL = 0), rates (n_cells + 1,
 * [0] unused); c_final (nspecies*n) optional. */
int lbib_adj_forward_ms(lbib_adj* adj, int steps, const double* tau_c, double kappa, double k_junction,
                        const double* K, const double* rates, double* loss, double* c_final);
/* Gradient of the last forward's loss with respect to every parameter of
 * lbib_adj_forward_ms (any output may be NULL). */
int lbib_adj_gradient_ms(lbib_adj* adj, double* d_tau_c, double* d_kappa, double* d_k_junction, double* d_K,
                         double* d_rates);

/* ---- measurement ------------------------------------------------------------
 * No counterpart in the reference (it has no profiling): HIP events recorded
 * on the context's stream around launches of the fused lattice kernel, every
 * on-th launch after the call (on = 1: every launch; 0 disables).  Each pair
 * of event packets costs the stream a few microseconds of idle time, so a
 * period above 1 perturbs the timed run less.
 * lbib_profile_read synchronises, returns the summed kernel time, the count
 * of timed launches and the algorithmic bytes of one launch (population read
 * + write at storage precision: (2*9 + 2*5*nspecies) * bytes * nx * ny), and
 * resets. */
int lbib_profile_enable(lbib_ctx* ctx, int on);
int lbib_profile_read(lbib_ctx* ctx, double* kernel_ms, int64_t* launches, double* bytes_per_launch);
/* Attainable-bandwidth yardstick: a plain 16-byte-vector copy of nbytes
 * (read nbytes + write nbytes) on the context's stream, reps timed launches
 * after one warm-up; returns the average launch time.  Allocates and frees two
 * nbytes buffers, so it is a set-up/measurement call, never a per-step one. */
int lbib_profile_copy(lbib_ctx* ctx, int64_t nbytes, int reps, double* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif /* LBIB_H */
