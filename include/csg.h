/* libcsg.so — C-ABI of the MI355X (gfx950) geodesic force path.
 *
 * Drop-in boundary for the mesh-space calls of curvedSpaceSim: every entry point below replaces a reference interface
 * (cited per function, paths relative to the reference tree) with a batched, device-backed equivalent. No C++ types
 * cross this boundary; arrays are plain host pointers (the call copies them to and from HBM) unless a function says
 * the data stays resident on the device. No function throws; each returns a csg_status and csg_last_error() describes
 * the last failure of a context. Calls on one context must not overlap (one context per host thread).
 *
 * Conventions (those of the reference):
 *   - a mesh position is a face index plus three barycentric weights in the order of the face's corners
 *     (inc/pointDataType.h:10-28; triangles are given counter-clockwise seen from outside);
 *   - tangents are unit vectors; the start tangent points from the source along the path, the end tangent points
 *     away from the source at the target (src/utility/meshUtilities.cpp:373-379);
 *   - with submeshing on, a target that the patch does not connect to the source gets distance 2*maximumDistance and
 *     tangents (0,0,1) (src/models/triangulatedMeshSpace.cpp:197-203).
 */
#ifndef CSG_H
#define CSG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csg_ctx csg_ctx;

typedef enum csg_status {
    CSG_OK = 0,
    CSG_ERR_ARGUMENT = -1,   /* bad argument (null pointer, size, index out of range) */
    CSG_ERR_DEVICE = -2,     /* HIP runtime error, or no gfx950 device */
    CSG_ERR_MESH = -3,       /* not a consistently oriented 2-manifold triangle mesh (triangulatedMeshSpace.cpp:51-63) */
    CSG_ERR_STATE = -4,      /* call order: no mesh, no particles, ... */
    CSG_ERR_WALK = -5,       /* a displacement walk failed as the reference throws (triangulatedMeshSpace.cpp:522-548,
                                meshUtilities.cpp:184-195); the per-particle status array names the case */
    CSG_ERR_CAPACITY = -6    /* an input exceeds a capacity of this build (documented per function) */
} csg_status;

/* pair potentials of the force computers (src/forces/): kind 0 harmonicRepulsion (p0 = k, p1 = sigma,
 * harmonicRepulsion.cpp:3-34), 1 gaussianRepulsion (p0 = alpha, p1 = sigma, gaussianRepulsion.cpp:3-12),
 * 2 truncated Gaussian (p0 = epsilon, p1 = s, p2 = cutoff; U = epsilon exp(-d^2 / 2 s^2) for d < cutoff, no reference
 * counterpart), 3 none. */
typedef struct csg_potential {
    int32_t kind;
    double p0, p1, p2;
} csg_potential;

/* per-pair flags */
#define CSG_PAIR_UNREACHED 1u /* no path inside the patch (or none shorter than rMax) */
#define CSG_PAIR_BEYOND 4u    /* a path exists but is longer than rMax */
#define CSG_PAIR_SENTINEL 16u /* unreachable pair reported with the reference sentinel */

/* --- context ------------------------------------------------------------------------------------------------------ */
csg_status csg_ctx_create(csg_ctx** out, int device);
void csg_ctx_destroy(csg_ctx* ctx);
const char* csg_last_error(const csg_ctx* ctx);
/* name of the device architecture the context runs on (e.g. "gfx950") */
const char* csg_device_name(const csg_ctx* ctx);

/* --- mesh: triangulatedMeshSpace::loadMeshFromFile (src/models/triangulatedMeshSpace.cpp:42-72) ----------------- */
csg_status csg_mesh_upload(csg_ctx* ctx, int32_t nV, const double* xyz, int32_t nF, const int32_t* tri);
/* getArea (triangulatedMeshSpace.cpp:153) and minVertexPosition / maxVertexPosition (:6-30, the box is seeded at the
 * origin as the reference does) */
csg_status csg_mesh_info(const csg_ctx* ctx, int32_t* nV, int32_t* nF, double* area, double* lo3, double* hi3);

/* --- triangulatedMeshSpace::useSubmeshingRoutines (src/models/triangulatedMeshSpace.h:60-66) --------------------- */
csg_status csg_set_submeshing(csg_ctx* ctx, int32_t enabled, double maximumDistance);

/* --- triangulatedMeshSpace::distance (src/models/triangulatedMeshSpace.cpp:155-238), batched over sources ---------
 * Source s has targets tgtOffsets[s] .. tgtOffsets[s+1]-1. distanceThreshold (nullable: 1e20 each) is the per-source
 * argument of the reference call; with submeshing the patch radius is min(maximumDistance, distanceThreshold[s]).
 * Outputs are per target: dist (nullable), startTan and endTan (3 each, nullable), flags (nullable).
 * Capacity: with submeshing, at most 2^20 targets per source; whole-mesh mode needs 2*nF hash slots per workgroup. */
csg_status csg_distance_batch(csg_ctx* ctx, int32_t nSrc, const int32_t* srcFace, const double* srcBary, const int32_t* tgtOffsets,
                              const int32_t* tgtFace, const double* tgtBary, const double* distanceThreshold, double* dist,
                              double* startTan, double* endTan, uint32_t* flags);

/* --- fused simpleModel::findNeighbors + force::computeForces (src/models/simpleModel.cpp:68-112,
 * src/forces/baseForce.cpp:12-44) over explicit neighbour lists of particles given by position -------------------
 * Particles: nP positions. Source s is particle srcIdx[s] (nullable: s) with candidate particles
 * tgtIdx[tgtOffsets[s] .. tgtOffsets[s+1]-1]; patchRadius (nullable: whole mesh) per source; windows whose distance
 * lower bound exceeds rMax are not propagated (pass a huge value for none; with a potential, rMax is raised to the
 * potential's range, and truncation is switched off when the sentinel distance lies inside that range, so that
 * forces and energies stay exact; only the reported distances of pairs beyond rMax change). Unreached pairs get the reference
 * sentinel when useSentinel != 0. forces: 3 per source, sum over its pairs of pairForce(startTangent, distance);
 * energy: per source sum of pair energies. Every output is nullable. */
csg_status csg_pair_forces(csg_ctx* ctx, int32_t nP, const int32_t* face, const double* bary, int32_t nSrc, const int32_t* srcIdx,
                           const int32_t* tgtOffsets, const int32_t* tgtIdx, const double* patchRadius, double rMax,
                           const csg_potential* pot, int32_t useSentinel, double sentinelDistance, double* dist, float* startTan,
                           float* endTan, uint32_t* flags, double* forces, double* energy);

/* --- csg_pair_forces for callers that cannot allocate (no reference counterpart) ----------------------------------
 * For code under HIP graph capture, tight serving loops and harnesses that forbid allocation: the arguments of
 * csg_pair_forces, but every array is a device pointer the caller owns (inputs and outputs), and the call allocates
 * nothing. Its scratch (patch lists, per-source status, spill queue, HBM tables for large patches) is a static device
 * array of 64 MiB that each GPU the library runs on holds until the process ends. Limits: at most 2048 sources per call
 * (CSG_ERR_CAPACITY beyond, before any work), patch radii required (no whole-mesh solves), and sources whose patches
 * exceed 1536 faces or 4096 targets make the call return CSG_ERR_CAPACITY after it waited for the stream (their outputs
 * are then not valid). Synchronous. Calls that overlap on one GPU race on the shared scratch. */
csg_status csg_pair_forces_noalloc(csg_ctx* ctx, int32_t nP, const int32_t* face, const double* bary, int32_t nSrc, const int32_t* srcIdx,
                                   const int32_t* tgtOffsets, const int32_t* tgtIdx, const double* patchRadius, double rMax,
                                   const csg_potential* pot, int32_t useSentinel, double sentinelDistance, double* dist, float* startTan,
                                   float* endTan, uint32_t* flags, double* forces, double* energy);

/* --- open meshes: the space class decides what a walk does at the mesh border ------------------------------------
 * CSG_BOUNDARY_CLOSED: triangulatedMeshSpace (a border edge or vertex is a walk error, triangulatedMeshSpace.cpp:522-548);
 * CSG_BOUNDARY_ABSORBING: absorbingOpenMeshSpace (src/models/absorbingOpenMeshSpace.cpp:2-50), the particle stops where
 * it meets the border; CSG_BOUNDARY_TANGENTIAL: tangentialOpenMeshSpace (src/models/tangentialOpenMeshSpace.cpp:3-64), it
 * slides along the border. Both follow openMeshSpace::transportParticleAndVectors (src/models/openMeshSpace.cpp:114-238)
 * and project transported vectors that point over the border. Applies to csg_transport_batch and to every walk of the
 * particle system; default CSG_BOUNDARY_CLOSED. */
#define CSG_BOUNDARY_CLOSED 0
#define CSG_BOUNDARY_ABSORBING 1
#define CSG_BOUNDARY_TANGENTIAL 2
csg_status csg_set_boundary(csg_ctx* ctx, int32_t mode);

/* --- triangulatedMeshSpace::transportParticleAndVectors (src/models/triangulatedMeshSpace.cpp:448-564), batched -----
 * n particles; face, bary (3 each) and disp (3 each) are in/out as in the reference; vecs holds nVec (0..2)
 * transported vectors per particle (3 doubles each), in/out. status (nullable) receives the number of face crossings
 * of each particle, or a negative code when its walk failed (that particle keeps its input state); the call then
 * returns CSG_ERR_WALK. */
csg_status csg_transport_batch(csg_ctx* ctx, int32_t n, int32_t* face, double* bary, double* disp, int32_t nVec, double* vecs,
                               int32_t* status);

/* --- meshPositionToEuclideanLocation (src/models/triangulatedMeshSpace.cpp:82-106) ------------------------------- */
csg_status csg_positions(csg_ctx* ctx, int32_t n, const int32_t* face, const double* bary, double* xyz);

/* --- device-resident particle system: the batched seam of simpleModel (src/models/simpleModel.h:25-111) ----------
 * Positions, velocities, masses and forces live in HBM; every step is a fixed sequence of launches on the context
 * stream with no host synchronisation. Errors inside a step (candidate capacity, an unsolved source, a failed walk)
 * are counted on the device (csg_system_status) and reported by the next synchronous call that checks them
 * (csg_system_compute_forces, csg_system_move, csg_system_sync): CSG_ERR_WALK when walks failed since the last check
 * (those particles kept their positions, as the reference throws at triangulatedMeshSpace.cpp:522-548), and
 * CSG_ERR_CAPACITY when a source lost candidates to the candidate capacity or got zero force because it exceeded the
 * step's geodesic tables. */
/* simpleModel::initializeSimpleModel / setParticlePositions (simpleModel.cpp:7-35, :318-326); vel and mass nullable
 * (zero velocities, unit masses) */
csg_status csg_system_init(csg_ctx* ctx, int32_t N, const int32_t* face, const double* bary, const double* vel, const double* mass);
/* force computer and neighbour structure: pot and its range (force::maximumInteractionRange, baseForce.h:57); cellList
 * 1 = cellListNeighborStructure (cellListNeighborStructure.cpp), 0 = all-to-all baseNeighborStructure; windows with a
 * distance bound above rMax are not propagated (use the potential's cutoff, or a huge value); maxPatchFaces sizes the
 * HBM tables for sources the LDS tables cannot hold (0: 1536 faces; at most 4096 targets per source with a cell
 * list). Uses the context's submeshing setting. */
csg_status csg_system_configure(csg_ctx* ctx, const csg_potential* pot, double range, int32_t cellList, double rMax, int32_t maxPatchFaces);
/* simulation::computeForces (simulation.cpp:61-70): findNeighbors + pair forces; forces (3N) and energy nullable.
 * Synchronous: a source that the step's fixed tables cannot hold is solved again with HBM tables that grow until it
 * fits (as csg_pair_forces does); CSG_ERR_CAPACITY only if one still does not fit or candidates overflowed. */
csg_status csg_system_compute_forces(csg_ctx* ctx, double* forces, double* energy);
/* simpleModel::moveParticles (simpleModel.cpp:44-66) with host displacements (3N); velocities transported if asked;
 * CSG_ERR_WALK if a walk failed */
csg_status csg_system_move(csg_ctx* ctx, const double* disp, int32_t transportVelocity);
/* nSteps of velocityVerletNVE::performUpdate (velocityVerletNVE.cpp:3-29), asynchronous: errors are reported by the
 * next csg_system_sync */
csg_status csg_system_nve(csg_ctx* ctx, double dt, int32_t nSteps);
/* nSteps of noseHooverNVT::performUpdate (noseHooverNVT.cpp:42-139) with chain length M, asynchronous; the chain is
 * initialised on the first call (setBathVariables, :28-36) and kept across calls with the same (T, tau, M) */
csg_status csg_system_nvt(csg_ctx* ctx, double dt, double T, double tau, int32_t M, int32_t nSteps);
/* waits for the context stream, then reports the errors counted since the last check (see above) */
csg_status csg_system_sync(csg_ctx* ctx);
/* copies state back (every pointer nullable): faces N, barycentric 3N, velocities 3N, forces 3N */
csg_status csg_system_get(csg_ctx* ctx, int32_t* face, double* bary, double* vel, double* forces);
/* asynchronous snapshot for trajectory writers (the state simpleModelDatabase::writeState records,
 * src/databases/simpleModelDatabase.cpp:63-120, taken without stalling the steps): csg_system_snapshot queues, behind
 * the steps queued so far, the Euclidean positions' computation and the copies of faces, barycentrics, positions,
 * velocities and forces into a pinned host buffer the system owns, and returns at once; steps queued afterwards run
 * while the host waits. csg_system_snapshot_read waits for those copies only and copies the snapshot into the caller's
 * arrays (every pointer nullable: faces N, barycentric 3N, positions 3N, velocities 3N, forces 3N). One snapshot is
 * held at a time: a later csg_system_snapshot replaces it. */
csg_status csg_system_snapshot(csg_ctx* ctx);
csg_status csg_system_snapshot_read(csg_ctx* ctx, int32_t* face, double* bary, double* xyz, double* vel, double* forces);
/* checkpoint and resume. simpleModelDatabase::readState (src/databases/simpleModelDatabase.cpp:122-158) restores
 * positions, velocities and forces: csg_system_set_state does the same from host arrays (every pointer nullable: that
 * field is kept). The reference saves no thermostat state, so a resumed Nose-Hoover run restarts its chain; here the
 * chain can be read and set as well, for a resume that continues the same trajectory. csg_system_thermostat: M (0 when
 * no NVT step ran since csg_system_init), T, tau, chain (4 (M + 1) doubles: position, velocity, acceleration, mass of each
 * entry; nullable) and the kinetic energy the chain last computed; csg_system_set_thermostat takes the same values, and
 * the next csg_system_nvt with the same (T, tau, M) continues from them (M = 0 clears the chain: the next NVT step
 * initialises a new one). */
csg_status csg_system_set_state(csg_ctx* ctx, const int32_t* face, const double* bary, const double* vel, const double* forces);
csg_status csg_system_thermostat(csg_ctx* ctx, int32_t* M, double* T, double* tau, double* chain, double* kineticEnergy);
csg_status csg_system_set_thermostat(csg_ctx* ctx, int32_t M, double T, double tau, const double* chain, double kineticEnergy);
/* counters since csg_system_init: pairs of the last force computation, candidate-capacity overflows, sources left
 * unsolved, failed walks; and the thermostat's kinetic energy (0 before any NVT step) */
csg_status csg_system_status(csg_ctx* ctx, int64_t* pairsLastForce, int32_t* candOverflow, int32_t* unsolved, int32_t* walkFailures,
                             double* kineticEnergy);
/* event timing of the step's launches (no reference counterpart; replaces the wall-clock profiler of
 * src/utility/profiler.h:16-59): when enabled, HIP events on the context stream bracket each launch group; ms[6] and
 * launches[6] receive the totals since the last query for {neighbour search, geodesic LDS kernel, geodesic HBM
 * kernel, walk kernel, patch kernel, learned screen} */
csg_status csg_system_profile(csg_ctx* ctx, int32_t enable);
csg_status csg_system_kernel_times(csg_ctx* ctx, double* ms, int32_t* launches);
/* work statistics of the last force computation: sums[7] and maxima[7] over sources of {patch faces, patch vertices,
 * FIFO steps, work items (windows and pseudo-source items), and the durations of patch construction, propagation and
 * epilogue in 10 ns ticks}; either pointer nullable */
csg_status csg_system_patch_stats(csg_ctx* ctx, int64_t* sums, int32_t* maxima);
/* sources handed from the LDS kernel to the HBM kernel since csg_system_init, by reason: reasons[6] = {patch faces,
 * patch vertices, live windows, targets, step limit, search frontier} */
csg_status csg_system_fallback_counts(csg_ctx* ctx, int32_t* reasons);
/* variants of the LDS geodesic kernel: 0 and 1: 256 and 512 lanes per source, two workgroups per CU; 2: 256 lanes, three
 * workgroups per CU with smaller tables and work-item ring; 3: 256 lanes, four workgroups per CU (40 KB of LDS each:
 * 320 patch vertices, 128 targets, a 280-record ring). Sources that do not fit a variant's tables are solved by the
 * HBM-table kernel. */
#define CSG_LDS_VARIANTS 4
/* autotuner (no reference counterpart): times the force computation of the current state with each LDS kernel variant
 * and keeps the fastest for later steps; chosen and ms[CSG_LDS_VARIANTS] (per force computation) nullable. The forces
 * of the state are recomputed, the particles are not moved. */
csg_status csg_system_autotune(csg_ctx* ctx, int32_t* chosen, double* ms);
/* graph mode (no reference counterpart): csg_system_nve / csg_system_nvt capture one step as a HIP graph on first use
 * for their parameters and replay it per step, which removes the per-launch host cost; ignored while event timing
 * (csg_system_profile) is on. Any configuration change drops the captured step. */
csg_status csg_system_use_graph(csg_ctx* ctx, int32_t enable);
/* sets the LDS kernel variant directly (0 .. CSG_LDS_VARIANTS - 1, as above) */
csg_status csg_system_set_lds_variant(csg_ctx* ctx, int32_t variant);
/* candidate lists of the last force computation (offsets N+1, indices up to cap), before any screening */
csg_status csg_system_pairs(csg_ctx* ctx, int32_t* offsets, int32_t* idx, int64_t cap);
/* independent particle groups in one system (no reference counterpart: many independent simulations stepped by the
 * same launches, each as the reference would run it alone). groupOf[i] in [0, nGroups) per particle; per group its
 * potential, interaction range (the cell list's candidate cutoff, at most the configured range, which sets the cell
 * size) and maximum distance (patch-radius cap and sentinel, as csg_set_submeshing's for a single system). Particles of
 * different groups are never candidates of each other, so groups on disjoint meshes of one uploaded union do not
 * interact. A group's range is also its truncation radius (raised as rMax is for a single system). Requires the cell
 * list. With the learned screen on, a pair's features and guard distance use its group's range. nGroups = 0 clears,
 * and so does csg_system_configure. */
csg_status csg_system_set_groups(csg_ctx* ctx, int32_t nGroups, const int32_t* groupOf, const csg_potential* pots, const double* ranges,
                                 const double* maximumDistances);
/* keep the geodesic distance of every candidate pair of later force computations (no reference counterpart; the
 * training data of the learned screen); csg_system_pair_distances copies those of the last one (screen off) */
csg_status csg_system_record_pairs(csg_ctx* ctx, int32_t enable);
csg_status csg_system_pair_distances(csg_ctx* ctx, double* dist, int64_t cap);

/* --- learned pair screen (no reference counterpart) ------------------------------------------------------------------
 * An MLP 16 -> 64 -> 64 -> 1 with ReLU scores how likely a candidate pair lies within the cutoff; it runs in BF16 on the
 * matrix cores with FP32 accumulation. Parameters: CSG_SCREEN_PARAMS floats, W1[16][64], b1[64], W2[64][64], b2[64],
 * W3[64], b3 (W1[i][j] multiplies feature i into hidden unit j). Features of a pair (i, j), with d the Euclidean
 * separation x_j - x_i, rc the potential's range, n the face normals, H and K the mean and Gaussian curvature
 * interpolated from the vertices, and h_i = d.n_i, h_j = d.n_j:
 *   100 (1 - |d|/rc), n_i.n_j, n_i (3), n_j (3), H_i rc, K_i rc^2, H_j rc, K_j rc^2, patch faces of i / 1000,
 *   100 h_i^2 / (6 |d| rc), 100 h_j^2 / (6 |d| rc), 10 h_i / |d|
 * (the last three estimate the arc-over-chord excess and the chord's slope). */
#define CSG_SCREEN_FEATURES 16
#define CSG_SCREEN_PARAMS 5313
/* logits of n feature rows */
csg_status csg_screen_forward(csg_ctx* ctx, const float* params, int32_t n, const float* features, float* logits);
/* gradient (CSG_SCREEN_PARAMS floats) and value of the mean binary cross-entropy with logits over n labelled rows */
csg_status csg_screen_gradient(csg_ctx* ctx, const float* params, int32_t n, const float* features, const float* labels, float* grad,
                               float* loss);
/* steps of Adam (beta 0.9, 0.999) on minibatches of one seeded shuffle of the rows, on the device; params in/out;
 * losses (nullable) receives each step's minibatch loss */
csg_status csg_screen_train(csg_ctx* ctx, float* params, int32_t n, const float* features, const float* labels, int32_t steps,
                            int32_t batch, float lr, uint32_t seed, float* losses);
/* features of the candidate pairs of the last force computation (cell list and submeshing required) */
csg_status csg_system_screen_features(csg_ctx* ctx, float* features, int64_t cap);
/* screen later force computations: pairs whose logit is below threshold and whose Euclidean separation is at least
 * 0.985 rc (rc: the range, or the pair's group range) are not solved (params NULL: screen off); closer pairs are kept
 * without being scored. Patches are still built from every candidate, so every kept pair gets the distance of the plain
 * computation. */
csg_status csg_system_set_screen(csg_ctx* ctx, const float* params, float threshold);
/* candidate pairs and kept pairs of the last force computation */
csg_status csg_system_screen_stats(csg_ctx* ctx, int64_t* candidates, int64_t* kept);

#ifdef __cplusplus
}
#endif
#endif
