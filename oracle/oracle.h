/* ORACLE — test infrastructure only.
 *
 * C interface of the FP64 CPU restatement of the reference's geodesic force path. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library, and only as the checker; the product never links it.
 *
 * Parity status of each function (how it is pinned independently of this code):
 *   orc_mesh_create     the reference's mesh checks: out-of-range indices, degenerate faces, non-manifold edges and
 *                       inconsistent orientation are rejected
 *   orc_geodesic        brute-force unfolding on tiny meshes, closed forms (plane, cube, tetrahedron), symmetry,
 *                       triangle inequality, d >= Euclidean, full mesh = patch for short pairs
 *   orc_patch           hand-worked cases of the reference rule on a small strip mesh and invariants
 *   orc_transport       flat-mesh exact translation, cube unfolding, speed preservation, reversibility; interior vertex
 *                       crossings worked by hand through the reference's steps (a flat grid and a cone apex, where the
 *                       heading is rotated); open-mesh border edges (absorbing stop point, tangential slide, vector
 *                       projection) against closed forms on flat meshes; a border vertex in both modes on exact dyadic
 *                       inputs worked by hand, including the tangential continuation
 *   orc_cell_candidates brute-force O(N^2) Euclidean test and the reference cell order on hand-built grids
 *   orc_pair_force      closed forms of the potentials
 *   orc_compute_forces, orc_move, with pyoracle.OracleSystem's integrators: a free particle moves on a straight line,
 *                       a head-on collision on a flat strip matches its closed-form trajectory, a cylinder run equals
 *                       its unrolled plane run, velocity Verlet's energy error falls as dt^2, and the Nose-Hoover
 *                       chain's extended energy is conserved (chains of 2 and 3)
 * Functions that only a CGAL build could pin (CGAL's rounding and its choice among tied geodesics) are reported as
 * "parity unpinned" where the tests use them.
 */
#ifndef CSG_ORACLE_H
#define CSG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh orc_mesh;

const char* orc_last_error(void);
orc_mesh* orc_mesh_create(int nV, const double* X, int nF, const int* F);
void orc_mesh_destroy(orc_mesh* m);
int orc_mesh_counts(const orc_mesh* m, int* nV, int* nF, int* nBorderVerts, int* nSaddle);
int orc_mesh_adjacency(const orc_mesh* m, int* adj, int* adjEdge);
int orc_mesh_angle_sums(const orc_mesh* m, double* out);

int orc_patch(const orc_mesh* m, int srcFace, const double* srcBary, int nT, const int* tgtFace, double radius, int* outFaces, int cap);
int orc_geodesic(const orc_mesh* m, const int* patchFaces, int nPatch, int srcFace, const double* srcBary, int nT, const int* tgtFace,
                 const double* tgtBary, double rMax, double* dist, double* startTan, double* endTan, int* flags, long long* stats);
int orc_distance(const orc_mesh* m, int submeshing, double maximumDistance, int srcFace, const double* srcBary, int nT, const int* tgtFace,
                 const double* tgtBary, double distanceThreshold, double* dist, double* startTan, double* endTan, int* flags);

/* walk behaviour at the mesh border: 0 closed surface (error), 1 absorbing, 2 tangential */
int orc_mesh_set_boundary(orc_mesh* m, int mode);
int orc_transport(const orc_mesh* m, int* face, double* bary, double* disp, int nVec, double* vecs);
int orc_transport_batch(const orc_mesh* m, int n, int* face, double* bary, double* disp, int nVec, double* vecs, int* status);
void orc_positions(const orc_mesh* m, int n, const int* face, const double* bary, double* out);

long long orc_cell_candidates(int N, const double* pos, const double* lo, const double* hi, double gridSize, double range, int* offsets,
                              int* idx, double* maxDist, long long cap);
double orc_pair_force(int kind, const double* params, double d, double* scaleOut);
int orc_compute_forces(const orc_mesh* m, int N, const int* face, const double* bary, int submeshing, double maximumDistance,
                       const int* offsets, const int* idx, const double* maxDist, int kind, const double* params, double* forces,
                       double* pairDist, double* pairTan, double* pairEnd, int* pairFlags, double* energy, int threads);
int orc_move(const orc_mesh* m, int N, int* face, double* bary, const double* disp, double* vel, int threads);

#ifdef __cplusplus
}
#endif
#endif
