/*
 * d3d.h -- C ABI of libd3d_hip.so, the MI355X (gfx950) collision hot path.
 *
 * Every entry point:
 *   - returns 0 on success and a nonzero D3D_E_* code on failure (it never
 *     throws across the ABI); d3d_last_error() then describes the failure
 *     (thread-local string);
 *   - takes caller-owned DEVICE pointers for every array argument (for
 *     example torch.Tensor.data_ptr() of a tensor on the GPU) unless the
 *     argument is documented as host memory;
 *   - enqueues work on `stream` (a hipStream_t passed as void*; NULL is the
 *     default stream) and does not synchronise unless documented.
 *
 * The reference (distance3d 0.9.0) is a Python library with no FFI; each
 * entry point below replaces the reference function cited beside it, batched
 * over many pairs.  The Python package distance3d_amd binds these with ctypes
 * and keeps the reference's call surface (gjk.gjk, epa.epa, colliders.*,
 * aabb_tree.AabbTree).
 */
#ifndef D3D_H
#define D3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D3D_ABI_VERSION 5

/* error codes */
#define D3D_OK 0
#define D3D_E_INVALID 1  /* bad argument (null pointer, size, shape index) */
#define D3D_E_HIP 2      /* a HIP runtime call failed */
#define D3D_E_CAPACITY 3 /* an output buffer was too small */

/* collider types (shape_i32[k*4 + 0]) */
#define D3D_HULL 0      /* ConvexHullVertices / Box-free vertex sets, MeshGraph */
#define D3D_SPHERE 1    /* Sphere */
#define D3D_CAPSULE 2   /* Capsule */
#define D3D_ELLIPSOID 3 /* Ellipsoid */
#define D3D_CYLINDER 4  /* Cylinder */
#define D3D_BOX 5       /* Box */
#define D3D_CONE 6      /* Cone */
#define D3D_DISK 7      /* Disk */
#define D3D_ELLIPSE 8   /* Ellipse */

/* GJK status (low byte = GjkState of gjk/_gjk_jolt.py:22-26) */
#define D3D_SEPARATED 0
#define D3D_INTERSECT 1
#define D3D_CLIPPED 3
#define D3D_ITER_CAP 4
#define D3D_SANITY_FAILED 0x100 /* bit set when |d.d - v_len_sq| >= sanity_check */

/*
 * A scene: the collider table and the vertex pools, resident in HBM.
 *   shape_f64[k*16 + 0..8]  R = collider2origin[:3,:3], row major
 *   shape_f64[k*16 + 9..11] t = collider2origin[:3,3]
 *   shape_f64[k*16 + 12..14] type parameters:
 *       HULL: p0 = max_i(|x_i|+|y_i|+|z_i|) over the local vertices,
 *             p1 = index of the hull's first cluster record, p2 = number of
 *             clusters (0: unclustered; otherwise ceil(n/D3D_CLUSTER_SIZE),
 *             and the hull's FP32-pool offset is a multiple of
 *             D3D_CLUSTER_SIZE)
 *       SPHERE: radius; CAPSULE: radius, height; CYLINDER: radius, length;
 *       CONE: radius, height; DISK: radius (R columns = plane basis, normal);
 *       ELLIPSE: radius0, radius1 (R columns 0,1 = axes); BOX: size[3];
 *       ELLIPSOID: radii[3]
 *   shape_f64[k*16 + 15]    margin (Margin collider), 0 otherwise
 *   shape_i32[k*4 + 0..3]   type, n_vertices, offset into vx/vy/vz,
 *                           offset into v64 (-1 when the FP32 values are exact)
 *   vx, vy, vz              FP32 vertex pool, structure of arrays, local frame
 *   v64                     FP64 vertex pool, (n, 3) row major, for hulls
 *                           whose vertices are not exactly representable in
 *                           FP32 (may be NULL when no shape uses it)
 *   vorig                   original index of each FP32-pool vertex within
 *                           its hull (NULL: pool order is the original order);
 *                           ties of the support argmax go to the smallest
 *   clusters                (n_clusters, 4) FP32 records {cx, cy, cz, r} of a
 *                           clustered hull: its vertices j*K .. j*K+K-1
 *                           (K = D3D_CLUSTER_SIZE, in pool order) lie within
 *                           the sphere, inflated so that the FP32 bound
 *                           c.d + r|d| of the sweep is never below an exact
 *                           vertex score (may be NULL when n_clusters == 0)
 *
 * d3d_cluster_scene builds a clustered copy of a scene on the host.
 * Alignment: shape_i32, vx, vy, vz, vorig and clusters start on 16-byte
 * boundaries, shape_f64 and v64 on 8-byte ones (device allocations are);
 * entry points taking a scene return D3D_E_INVALID otherwise.
 */
#define D3D_CLUSTER_SIZE 16

typedef struct d3d_scene {
    const double *shape_f64;
    const int32_t *shape_i32;
    const float *vx;
    const float *vy;
    const float *vz;
    const double *v64;
    int64_t n_shapes;
    int64_t n_vertices;
    int64_t n_vertices64;
    const int32_t *vorig;
    const float *clusters;
    int64_t n_clusters;
} d3d_scene;

/* gjk_distance_jolt keyword arguments (gjk/_gjk_jolt.py:138-139) plus the
 * iteration cap the reference loop lacks and the launch tuning.
 *
 * Work classes: d3d_gjk_distance_batch sorts pairs by the total number of
 * hull vertices n = n_A + n_B (0 for analytic colliders) and by whether both
 * colliders are hulls:
 *   size class 0: n = 0, 1: n <= 32, 2: <= 128, 3: <= 512, 4: <= 1024,
 *   5: <= 2048, 6: <= 4096, 7: > 4096;
 *   work class = size class, plus 8 when both colliders are hulls (D3D_NUM_CLASSES).
 * lanes_per_pair[c] lanes (a power of two <= 64; 0 = built-in default
 * 1,4,16,8,16,16,32,64 for one hull and 1,8,16,8,64,16,64,64 for two) work on
 * each pair of class c (class 0: 1, 2 or 4; larger values use 4).  For hull classes with 4 or more lanes, a pair whose hull
 * vertices fit the group's LDS slot (44, 85, 216, 456 or 1008 vertices with
 * 4, 8, 16, 32 or 64 lanes) has them copied there once and is swept from LDS;
 * when both colliders are hulls the two halves of the group sweep A and B
 * concurrently.
 * Launch shape (0 = built-in default): blocks_per_cu persistent GJK blocks of
 * four wavefronts per compute unit (1..3, default 3); claim_factor: a
 * wavefront claims claim_factor * 64 / G pairs of a class at a time (1..8,
 * default 2; fewer when the class is small); epa_blocks_per_cu persistent EPA
 * wavefronts per compute unit in the fused pipeline (1..16, default 8). */
#define D3D_NUM_CLASSES 16
typedef struct d3d_gjk_params {
    double tolerance;            /* 1e-10 */
    double max_distance_squared; /* 1e5 */
    double sanity_check;         /* 1e-8 */
    int32_t max_iterations;      /* 128 */
    int32_t hull_serial_max;     /* intersection / support kernels: hulls with
                                    <= this many vertices are swept by their
                                    own lane, larger ones by the wavefront */
    int32_t lanes_per_pair[D3D_NUM_CLASSES];
    int32_t blocks_per_cu;
    int32_t claim_factor;
    int32_t epa_blocks_per_cu;
} d3d_gjk_params;

/* Library version / device setup. */
int d3d_abi_version(void);
const char *d3d_last_error(void);
/* Selects `device` for the calling thread and checks it is a gfx950 GPU. */
int d3d_init(int device);

/* ---- scene layout for pruned support sweeps (HOST memory, no GPU needed) --
 * Replaces nothing in the reference: it reorders each ConvexHullVertices
 * vertex set (colliders.py:109-144) into spatial clusters so the argmax of
 * support_function (colliders.py:131-132) can skip clusters whose bound lies
 * below the running maximum; results are unchanged bit for bit.
 *
 * d3d_cluster_plan: output sizes for d3d_cluster_scene.  Hulls with at least
 * min_vertices vertices (min_vertices > 0) are clustered.
 * d3d_cluster_scene: writes the clustered scene.  Inputs are a host scene in
 * the layout above (unclustered: p1 = p2 = 0, vorig NULL); outputs are host
 * arrays sized by the plan: out_f64 (n_shapes,16), out_i32 (n_shapes,4),
 * ovx/ovy/ovz (n_vertices), ov64 (n_vertices64,3), vorig (n_vertices),
 * clusters (n_clusters,4).  Hull vertex ranges are re-laid out (a clustered
 * hull starts at a multiple of D3D_CLUSTER_SIZE); records sharing one vertex
 * range keep sharing it.  n_threads <= 0 uses all hardware threads. */
int d3d_cluster_plan(const int32_t *shape_i32, int64_t n_shapes, int32_t min_vertices,
                     int64_t *n_vertices, int64_t *n_vertices64, int64_t *n_clusters);
int d3d_cluster_scene(const double *shape_f64, const int32_t *shape_i32, int64_t n_shapes,
                      const float *vx, const float *vy, const float *vz, const double *v64,
                      int64_t n_vertices, int64_t n_vertices64, int32_t min_vertices,
                      int32_t n_threads, double *out_f64, int32_t *out_i32, float *ovx,
                      float *ovy, float *ovz, double *ov64, int32_t *vorig, float *clusters);

/*
 * Support points, collider.support_function(d) (colliders.py:131-132,
 * 272-274, 326-328, 374-376, 426-428, 479-481, 533-535, 590-592, 629-631;
 * geometry.py:160-454).
 *   shape (n) int32, dirs (n,3) f64 world directions -> out (n,3) f64.
 */
int d3d_support_batch(const d3d_scene *scene, const int32_t *shape, const double *dirs,
                      int64_t n, double *out, void *stream);

/*
 * Axis-aligned bounding boxes, collider.aabb() (colliders.py:140-141,
 * 180-181, 281-282, 335-337, 383-384, 435-437, 489-490, 543-544, 599-600,
 * 639-643; containment.py:6-229).
 *   shape (n) int32 -> out (n,3,2) f64 [axis][min,max].
 */
int d3d_aabb_batch(const d3d_scene *scene, const int32_t *shape, int64_t n, double *out,
                   void *stream);

/*
 * Batched GJK distance, gjk.gjk / gjk_distance_jolt (gjk/_gjk_jolt.py:138-221).
 *   pair_a, pair_b (n) int32 shape indices.
 *   dist (n) f64: distance, DBL_MAX when clipped (the reference's MAX_FLOAT).
 *   cp_a, cp_b (n,3) f64: closest points (NaN when clipped).
 *   status (n) int32: D3D_SEPARATED / INTERSECT / CLIPPED / ITER_CAP, with
 *       D3D_SANITY_FAILED or-ed in when the reference's sanity assert
 *       (gjk/_gjk_jolt.py:215-216) would fire.
 *   iterations (n) int32, optional (NULL).
 *   simplex (n,3,4,3) f64, optional: final Y, P, Q rows (the reference's
 *       returned simplex is the Y block); n_points (n) int32, optional.
 * Uses a workspace the library keeps per (device, stream), allocated with
 * hipMalloc on first use and grown on demand.
 */
int d3d_gjk_distance_batch(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                           int64_t n, const d3d_gjk_params *params, double *dist, double *cp_a,
                           double *cp_b, int32_t *status, int32_t *iterations, double *simplex,
                           int32_t *n_points, void *stream);

/*
 * Stage timing of the GJK pipelines (d3d_gjk_distance_batch*,
 * d3d_gjk_epa_batch*); no counterpart in the reference.  While enabled, every
 * call records HIP events on its stream before the counting sort, before the
 * GJK kernel, after it and after the EPA kernel.  d3d_stage_timing(1) starts
 * (dropping what was recorded), d3d_stage_timing(0) stops.
 * d3d_stage_timing_read waits for the recorded events, writes the summed
 * milliseconds of the three stages (counting sort, GJK kernel, EPA kernel; 0
 * without EPA) to ms_out[3] and the number of calls to calls_out, and forgets
 * them.  Meant for measurement: calls under graph capture must not run with it
 * enabled.
 */
int d3d_stage_timing(int enable);
int d3d_stage_timing_read(double *ms_out, int64_t *calls_out);

/* Bytes of device workspace d3d_gjk_distance_batch_ws needs for n pairs. */
int64_t d3d_gjk_workspace_bytes(int64_t n);

/*
 * d3d_gjk_distance_batch with a caller-owned device workspace of at least
 * d3d_gjk_workspace_bytes(n) bytes (no allocation inside; calls that overlap
 * in time must use distinct workspaces).  Same outputs as
 * d3d_gjk_distance_batch (gjk/_gjk_jolt.py:138-221).
 */
int d3d_gjk_distance_batch_ws(const d3d_scene *scene, const int32_t *pair_a,
                              const int32_t *pair_b, int64_t n, const d3d_gjk_params *params,
                              double *dist, double *cp_a, double *cp_b, int32_t *status,
                              int32_t *iterations, double *simplex, int32_t *n_points,
                              void *workspace, int64_t workspace_bytes, void *stream);

/*
 * Batched GJK intersection, gjk.gjk_intersection / gjk_intersection_jolt
 * (gjk/_gjk_jolt.py:29-135).
 *   hit (n) uint8: 1 intersecting, 0 separated.
 *   status (n) int32, optional: D3D_SEPARATED / INTERSECT / ITER_CAP.
 */
int d3d_gjk_intersection_batch(const d3d_scene *scene, const int32_t *pair_a,
                               const int32_t *pair_b, int64_t n, double tolerance,
                               int32_t max_iterations, uint8_t *hit, int32_t *status,
                               void *stream);

/*
 * Batched EPA penetration depth, epa.epa (epa.py:9-78), seeded from the GJK
 * simplices of d3d_gjk_distance_batch (simplex (n,3,4,3) f64 = Y, P, Q rows,
 * n_points (n) int32) of intersecting pairs.  One wavefront per pair, polytope
 * in LDS.
 *   max_iterations >= 1 (epa.py's max_iter), max_faces in [4, 128],
 *   epsilon: convergence tolerance (1e-8).
 *   out (n,13) f64: depth, normal[3], mtv[3] = normal * depth, contact point
 *       on A [3], contact point on B [3] (cp_a - cp_b = mtv when converged).
 *   status (n) int32: 1 converged, 0 max_iterations reached, 2 face cap (or
 *       the polytope's 72 vertex rows used up: 128 faces need about 66),
 *       3 degenerate seed (flat Minkowski difference), 0x200 bad input.
 *   iterations (n) int32, optional.
 *   faces (n, max_faces, 4, 3) f64, optional: the final polytope,
 *       Polytope.get_all_faces (epa.py:70,78): per face its three vertices
 *       (Minkowski points) and its outward unit normal; n_faces (n) int32,
 *       optional: the number of faces written per pair (rows past it are
 *       not touched).
 */
int d3d_epa_batch(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                  int64_t n, const double *simplex, const int32_t *n_points,
                  int32_t max_iterations, int32_t max_faces, double epsilon, double *out,
                  int32_t *status, int32_t *iterations, double *faces, int32_t *n_faces,
                  void *stream);

/*
 * GJK then EPA in one device-side pipeline (gjk.gjk followed by epa.epa on
 * the intersecting pairs, gjk/_gjk_jolt.py:138-221 and epa.py:9-78): GJK as
 * d3d_gjk_distance_batch, intersecting pairs compacted on the device, EPA on
 * them by a persistent kernel; no host synchronisation.
 *   epa_out (n,13) f64 as d3d_epa_batch; epa_status (n) int32 as
 *   d3d_epa_batch, or -1 for pairs that do not intersect (zero outputs).
 *   workspace: caller-owned, >= d3d_gjk_epa_workspace_bytes(n) bytes.
 */
int64_t d3d_gjk_epa_workspace_bytes(int64_t n);
int d3d_gjk_epa_batch_ws(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                         int64_t n, const d3d_gjk_params *params, int32_t epa_max_iterations,
                         int32_t epa_max_faces, double epa_epsilon, double *dist, double *cp_a,
                         double *cp_b, int32_t *status, int32_t *iterations, double *epa_out,
                         int32_t *epa_status, void *workspace, int64_t workspace_bytes,
                         void *stream);

/*
 * d3d_gjk_epa_batch_ws with no workspace argument, allocating nothing: no
 * hipMalloc, hipFree or allocating library call.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *   The workspace is an array inside the library of
 *   d3d_gjk_epa_workspace_bytes(D3D_NOALLOC_MAX_PAIRS) bytes (about 5 MB),
 *   one copy per GPU, held until the process ends.  n above
 *   D3D_NOALLOC_MAX_PAIRS is rejected with D3D_E_INVALID.
 *   Every call on one GPU works in the same array and it keeps the last
 *   call's contents: calls that overlap in time on one GPU (other streams,
 *   other host threads) race and must be serialised by the caller.
 *   Outputs as d3d_gjk_epa_batch_ws (gjk/_gjk_jolt.py:138-221, epa.py:9-78).
 */
#define D3D_NOALLOC_MAX_PAIRS 16384
int d3d_gjk_epa_batch_noalloc(const d3d_scene *scene, const int32_t *pair_a,
                              const int32_t *pair_b, int64_t n, const d3d_gjk_params *params,
                              int32_t epa_max_iterations, int32_t epa_max_faces,
                              double epa_epsilon, double *dist, double *cp_a, double *cp_b,
                              int32_t *status, int32_t *iterations, double *epa_out,
                              int32_t *epa_status, void *stream);

/*
 * Pose-updated pair instances (collider.update_pose per pair,
 * colliders.py:174-178, 276-279, 330-333, ...): writes a collider table of
 * 2n records -- record 2i = scene collider pair_a[i] with pose row
 * pose_a[i], record 2i+1 = pair_b[i] with pose row pose_b[i] (12 f64 per
 * row: R row-major, t; NULL keeps the scene pose) -- into inst_f64 (2n,16) /
 * inst_i32 (2n,4), and, when inst_a/inst_b are given, the pair lists 2i /
 * 2i+1.  With pose_idx_a / pose_idx_b (n int32) the rows are
 * pose_a[pose_idx_a[i]] of a table of n_poses rows (a pose table kept on the
 * device).  The instances keep the scene's vertex offsets: a d3d_scene made
 * of this table and the scene's vertex pools evaluates the posed pairs.
 */
int d3d_instance_pairs(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                       const double *pose_a, const double *pose_b, const int32_t *pose_idx_a,
                       const int32_t *pose_idx_b, int64_t n_poses, int64_t n, double *inst_f64,
                       int32_t *inst_i32, int32_t *inst_a, int32_t *inst_b, void *stream);

/*
 * Minkowski portal refinement for n pairs (mpr.py:20-49 mpr_intersection,
 * 52-109 mpr_penetration, with _discover_portal / _refine_portal /
 * _find_penetration_info, :120-393).  centers: (n_shapes, 3) f64, the
 * collider centres (collider.center()) indexed by pair_a / pair_b.
 * status[i]: 1 intersecting, 0 separated, 0x100 the portal refinement gave
 * up after max_iterations expansions (the reference loops without bound),
 * 0x200 invalid collider record.  With penetration != 0, out[7 i ..] =
 * depth, penetration direction (3), contact position (3) for intersecting
 * pairs (NaN otherwise); with penetration == 0 only status is written
 * (mpr_intersection) and out may be NULL.
 */
int d3d_mpr_batch(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                  const double *centers, int64_t n, double tolerance, int32_t max_iterations,
                  int32_t penetration, int32_t *status, double *out, void *stream);

/*
 * Hydroelastic contact of m tetrahedron pairs (hydroelastic_contact/
 * _tetrahedron_intersection.py:7-85 intersect_tetrahedron_pairs, 89-139
 * intersect_tetrahedron_pair and the contact plane / polygon routines it
 * calls, _halfplanes.py:8-64).  Body k has n_k tetrahedra: points T_k
 * (n_k,4,3) f64, pressure potentials e_k (n_k,4), barycentric transforms X_k
 * (n_k,4,4) (barycentric_transforms, _barycentric_transform.py:4-9); pairs
 * (m,2) int32 index (body 1, body 2).  Per pair: status (bit 0 the
 * tetrahedra intersect, 0x10 the two potential fields are the same plane
 * family -- the reference's same-tetrahedron case --, 0x20 the contact plane
 * cuts both tetrahedra (a polygon was computed), 0x100 more half-plane
 * intersection points than the reference reserves, where it asserts), the
 * contact plane in Hesse normal form planes (m,4), and the contact polygon
 * in counter-clockwise order n_vertices (m), polygons (m,24,3) (NaN past
 * n_vertices).  planes and polygons may be NULL (status and n_vertices
 * only, e.g. a first pass that finds the intersecting pairs of a large
 * broad-phase list).  workspace: d3d_tet_pairs_workspace_bytes(m) device
 * bytes (at most 100 MB whatever m).
 */
int64_t d3d_tet_pairs_workspace_bytes(int64_t m);
int d3d_tet_pairs_batch(const double *T1, const double *e1, const double *X1, const double *T2,
                        const double *e2, const double *X2, const int32_t *pairs, int64_t m,
                        double youngs_modulus1, double youngs_modulus2, int32_t *status,
                        double *planes, int32_t *n_vertices, double *polygons, void *workspace,
                        int64_t workspace_bytes, void *stream);

/*
 * The angle the contact-polygon ordering sorts by (order_points,
 * hydroelastic_contact/_tetrahedron_intersection.py:300-318, np.arctan2):
 * out[k] = atan2(y[k], x[k]) rounded correctly to FP64 (C99 special cases).
 * np.arctan2's last bit varies with the host's numpy build and C library;
 * this is the one fixed choice, shared with the pair kernel.
 */
int d3d_polygon_angles(const double *y, const double *x, int64_t n, double *out, void *stream);

/*
 * intersect_halfplanes (hydroelastic_contact/_halfplanes.py:29-64): the
 * pairwise intersections of n half-planes (rows p.x, p.y, pq.x, pq.y, f64)
 * that lie inside all the others, in the reference's (i, j > i) order, to
 * points (3n, 2); *count (device int32) = their number, or -1 where the
 * reference asserts (more points than it reserves).  One wavefront does
 * O(n^3) work, so n is limited to D3D_MAX_HALFPLANES (the reference calls it
 * with at most 8 half-planes per tetrahedron pair).
 */
#define D3D_MAX_HALFPLANES 1024
int d3d_intersect_halfplanes(const double *halfplanes, int32_t n, double *points, int32_t *count,
                             void *stream);

/*
 * Per-tetrahedron preparation of a tetrahedral mesh (T (n,4,3) f64): its
 * AABBs (n,3,2) [axis][min,max] (tetrahedral_mesh_aabbs,
 * hydroelastic_contact/_mesh_processing.py:4-20) and barycentric transforms
 * X (n,4,4), the inverse of [T^T; 1 1 1 1] by LU with partial pivoting
 * (barycentric_transforms, _barycentric_transform.py:4-9, which takes the
 * pseudo-inverse: the same matrix for a non-degenerate tetrahedron, equal
 * to rounding); flags (n) = 1 where a pivot is zero (degenerate
 * tetrahedron).  Any of aabbs, X, flags may be NULL (not computed).
 */
int d3d_tet_mesh_prep(const double *T, int64_t n, double *aabbs, double *X, int32_t *flags,
                      void *stream);

/*
 * Contact forces of tetrahedron pairs (compute_contact_force,
 * hydroelastic_contact/_forces.py:52-111, batched as contact_surface_forces
 * :6-24 does): for each pair whose status (from d3d_tet_pairs_batch) has bit
 * 0, from the first body's tetrahedron T1[pairs[k,0]], potentials
 * e1[pairs[k,0]] and Young's modulus and the pair's plane and polygon
 * (planes, n_vertices, polygons as d3d_tet_pairs_batch writes them): out
 * (m,7) = polygon centre (3), force vector (3), area; zero rows for the other
 * pairs.
 */
int d3d_tet_contact_forces(const double *T1, const double *e1, const int32_t *pairs,
                           const int32_t *status, const double *planes,
                           const int32_t *n_vertices, const double *polygons, int64_t m,
                           double youngs_modulus1, double *out, void *stream);

/*
 * Broad phase: linear BVH over n AABBs and overlap queries
 * (aabb_tree.py:15-191 AabbTree, 381-403 query_overlap, 465-527
 * all_aabbs_overlap / aabb_overlap).
 *
 * d3d_bvh_bytes(n): size of the caller-owned device buffer holding a BVH
 * over n boxes.  d3d_bvh_build(aabbs (n,3,2) f64 [axis][min,max], n, bvh,
 * bytes, stream) builds it (the aabbs must stay alive and unchanged while
 * the BVH is queried).  d3d_bvh_query(bvh, n, aabbs, queries (m,3,2) f64, m,
 * pairs (cap,2) int32, cap, n_pairs, stream) writes every (query index,
 * box index) with aabb_overlap(box, query) true -- closed intervals on all
 * three axes, FP64 -- in arbitrary order, and the total count to the device
 * uint64 *n_pairs (pairs beyond cap are counted but not written).
 */
int64_t d3d_bvh_bytes(int64_t n);
int d3d_bvh_build(const double *aabbs, int64_t n, void *bvh, int64_t bvh_bytes, void *stream);
int d3d_bvh_query(const void *bvh, int64_t n, const double *aabbs, const double *queries,
                  int64_t m, int32_t *pairs, int64_t cap, uint64_t *n_pairs, void *stream);

/*
 * all_aabbs_overlap (aabb_tree.py:465-500) by brute force: every (i, j) with
 * aabb_overlap(a[i], b[j]) -- a (n,3,2), b (m,3,2) f64 --, written to pairs
 * (cap,2) int32 in the reference's row-major order (i ascending, then j
 * ascending), the total count to the device uint64 *n_pairs (pairs beyond
 * cap are counted, not written).  Tests all n*m pairs: for small or densely
 * overlapping sets, where it is faster than a BVH query and needs no sort.
 * workspace: d3d_overlap_dense_workspace_bytes(n) device bytes.
 */
int64_t d3d_overlap_dense_workspace_bytes(int64_t n);
int d3d_overlap_dense(const double *a, int64_t n, const double *b, int64_t m, int32_t *pairs,
                      int64_t cap, uint64_t *n_pairs, void *workspace, int64_t workspace_bytes,
                      void *stream);

/*
 * d3d_bvh_refit(aabbs (n,3,2) f64, n, bvh, bytes, stream): new boxes for the
 * leaves of a BVH that d3d_bvh_build built over n boxes (box i keeps leaf
 * slot i), keeping its topology: the node boxes are recomputed bottom-up and
 * queries stay exact (a node box only has to contain its leaves' boxes).
 * `aabbs` replaces the array the build was given.  Replaces the full
 * rebuild by serial insertion on every pose update in
 * BoundingVolumeHierarchy.update_collider_poses (broad_phase.py:144-151).
 */
int d3d_bvh_refit(const double *aabbs, int64_t n, void *bvh, int64_t bvh_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* D3D_H */
