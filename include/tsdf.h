/*
 * tsdf.h -- C ABI of libtsdf_hip.so, the MI355X (gfx950) voxel-hashed TSDF
 * engine: block allocation into a spatial hash, projective integrate /
 * deintegrate, and the voxel-centre raycast.
 *
 * The ABI replaces the member functions of the reference class GPUVoxelGrid
 * (src/VoxelGrid/GPUVoxelGrid.h:40-120).  Each entry point names the
 * reference interface it stands in for.  Conventions:
 *   - plain pointers and sizes only; `stream` is a hipStream_t passed as
 *     void* (NULL = the null stream of the grid's device);
 *   - image inputs are DEVICE pointers, row-major [H][W] uint16 depth in mm
 *     (0 = no measurement) and optional [H][W][3] uint8 colour;
 *   - poses are HOST float[16], row-major camera->world, rigid (the last
 *     row must be 0 0 0 1, GPUVoxelGrid.cpp:52-57 / integrate.cpp:147-159);
 *   - calls are asynchronous on `stream`; tsdf_sync() / tsdf_stats() wait and
 *     report deferred device-side errors (pool overflow);
 *   - the caller owns its inputs, the handle owns the hash table and the
 *     brick pool; one handle is used by one thread at a time (the reference
 *     serialises callers with tsdfLock, integrate.h:23);
 *   - no exception or abort crosses the ABI: every failure is a status.
 */
#ifndef SUPERSCANNER_TSDF_H
#define SUPERSCANNER_TSDF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tsdf_grid tsdf_grid;

typedef enum {
    TSDF_OK = 0,
    TSDF_BAD_ARG = 1,        /* NULL handle/pointer, bad size or parameter          */
    TSDF_BAD_DIMS = 2,       /* image shape differs from the grid's (H, W); the     */
                             /* reference only prints here, GPUVoxelGrid.cpp:46-49   */
    TSDF_MAP_FULL = 3,       /* brick pool or hash table overflowed (the reference  */
                             /* leaves this unchecked, GPUVoxelGrid.cpp:497)        */
    TSDF_HIP_ERROR = 4,      /* a HIP runtime call failed; see tsdf_last_error()    */
    TSDF_NO_INTRINSICS = 5,  /* integrate/raycast before tsdf_set_intrinsics        */
    TSDF_BAD_POSE = 6        /* non-finite or non-rigid pose                        */
} tsdf_status;

enum {
    TSDF_STORAGE_F32 = 0,    /* fp32 sdf + fp32 weight, 8 B/voxel                    */
    TSDF_STORAGE_F16 = 1     /* fp16 sdf + fp16 weight, 4 B/voxel, fp32 arithmetic  */
};

enum {
    TSDF_RAY_UB_VOXEL_INDEX = 1, /* voxel index -1 or 8 from float rounding         */
                                 /* (GPUVoxelGrid.h:192-221): reference reads out of */
                                 /* the brick; here the pixel is zeroed and flagged  */
    TSDF_RAY_UB_MISSING_OUT = 2  /* zero crossing whose outer voxel lies in a block  */
                                 /* that is not in the map (GPUVoxelGrid.cpp:669-672)*/
};

typedef struct {
    float voxel_length;     /* vl in metres (GPUVoxelGrid ctor `voxlength`)          */
    float truncation;       /* d_trunc in metres (ctor `trunc`)                      */
    int32_t height;         /* resi (ctor `resi_`)                                   */
    int32_t width;          /* resj (ctor `resj_`)                                   */
    uint32_t n_buckets;     /* hash slots, power of two                              */
    uint32_t n_bricks;      /* brick pool capacity (reference: c / 512,              */
                            /* GPUVoxelGrid.cpp:11,27)                               */
    int32_t storage;        /* TSDF_STORAGE_*                                        */
    int32_t with_color;     /* keep a running-average rgb per voxel                  */
    float max_weight;       /* 0 = uncapped, as the reference (GPUVoxelGrid.h:62);   */
                            /* fp16 storage caps at 2048 (refuses more), the        */
                            /* largest weight fp16 counts exactly                   */
    float min_depth;        /* rmin, 0.2 m (GPUVoxelGrid.h:56)                       */
    float max_ray_length;   /* d_max, 10 m (GPUVoxelGrid.h:55)                       */
    int32_t device;         /* HIP device ordinal                                    */
} tsdf_params;

typedef struct {
    uint64_t n_bricks;      /* bricks allocated (voxelMap2.size())                   */
    uint64_t capacity;      /* pool capacity (maxVoxelBlocks)                        */
    uint64_t overflow;      /* allocations refused since the last reset              */
    uint64_t last_touched;  /* bricks in the last integrate call's touched set       */
    uint64_t total_touched; /* sum of touched-set sizes since reset                  */
    uint64_t total_new;     /* bricks inserted since reset                           */
    uint64_t total_updated; /* voxel updates since reset                             */
    uint64_t total_valid_px;/* pixels with camera z > rmin since reset               */
    uint64_t ray_hits;      /* raycast pixels with a zero crossing, since reset      */
    uint64_t ray_ub;        /* raycast pixels flagged as reference UB, since reset   */
    uint64_t ray_samples;   /* raycast voxel reads since reset                       */
    uint64_t ray_probes;    /* raycast hash lookups since reset                      */
} tsdf_stats_out;

/* Kernel kinds timed by tsdf_profile (no reference counterpart: the
 * reference names its kernels for Kokkos-Tools, kernelFilters.lst). */
enum {
    TSDF_K_RESET = 0,           /* table / counter reset                         */
    TSDF_K_COLLECT = 1,         /* block allocation ("Integration Collection")   */
    TSDF_K_INTEGRATE = 2,       /* brick update ("Integration Core Kernel")      */
    TSDF_K_COLLECT_BATCH = 3,   /* allocation for a batch of ops                 */
    TSDF_K_INTEGRATE_BATCH = 4, /* brick update for a batch of ops               */
    TSDF_K_RAYCAST = 5,         /* "Raycasting Kernel"                           */
    TSDF_N_KERNEL_KINDS = 6
};

typedef struct {
    double ms[TSDF_N_KERNEL_KINDS];         /* summed device time per kind   */
    uint64_t launches[TSDF_N_KERNEL_KINDS]; /* launches timed per kind       */
} tsdf_kernel_times;

/* Launch shapes of the persistent kernels (no reference counterpart; picked
 * per device and workload by superscanner_amd.autotune). */
typedef struct {
    int32_t raycast_xcd_map;      /* 1: each XCD takes one band of the image    */
    int32_t raycast_skip;         /* empty-space skip of the raycast (results   */
                                  /* do not depend on it): 0 none, 1 distance   */
                                  /* field over superblocks, 2 depth bounds of  */
                                  /* 8 x 8-pixel tiles from the projected bricks */
    int32_t integrate_wg_per_cu;  /* brick-update workgroups per CU            */
    int32_t batch_ops_per_pass;   /* ops per collection + update pass of a     */
                                  /* batch, 1..256                              */
    int32_t batch_graph;          /* 1: a batch of <= 256 ops runs as one       */
                                  /* captured graph (staging copy + passes)     */
    int32_t raycast_order;        /* order the raycast's tiles are dispatched   */
                                  /* in (results do not depend on it): 0 the    */
                                  /* hardware's, 1 reversed, 2 the tiles that   */
                                  /* took longest in the previous raycast first */
    int32_t raycast_budget;       /* march rounds of a ray before it is handed  */
                                  /* to a second kernel that carries each ray   */
                                  /* with 16 lanes (results do not depend on    */
                                  /* it); 0 = never, at most 4096               */
} tsdf_tuning;

/* Camera of a render view (the differentiable render has its own resolution
 * and intrinsics; the reference convention ci = height - cy, cj = cx holds). */
typedef struct {
    int32_t height, width;
    float fx, fy, cx, cy;
} tsdf_view;

/* Fills the defaults of the reference: d_max 10 m, rmin 0.2 m, no weight
 * cap, fp32 storage, no colour (GPUVoxelGrid.h:55-62). */
void tsdf_default_params(tsdf_params *p);

/* GPUVoxelGrid::GPUVoxelGrid(c, trunc, voxlength, resi, resj),
 * GPUVoxelGrid.cpp:10-30.  Allocates the table and pool on p->device. */
tsdf_status tsdf_create(const tsdf_params *p, tsdf_grid **out);

/* The parameters the grid runs with (the ones given to tsdf_create, with
 * the fp16 weight cap filled in); the reference's public members
 * d_trunc, voxlength, resi, resj, maxVoxelBlocks (GPUVoxelGrid.h:50-62). */
tsdf_status tsdf_get_params(const tsdf_grid *g, tsdf_params *p);

/* Releases everything the handle owns (reference: destructor of the Views). */
void tsdf_destroy(tsdf_grid *g);

/* Empties the map (all bricks freed, counters zeroed); no reference
 * counterpart beyond constructing a new GPUVoxelGrid. */
tsdf_status tsdf_reset(tsdf_grid *g, void *stream);

/* GPUVoxelGrid::setIntrinsic(space, PinholeCameraIntrinsic),
 * GPUVoxelGrid.cpp:253-266: stores fx, fy and the reference's
 * ci = H - cy, cj = cx convention. */
tsdf_status tsdf_set_intrinsics(tsdf_grid *g, float fx, float fy, float cx, float cy);

/* GPUVoxelGrid::integrate (sign = +1) and GPUVoxelGrid::deIntegrate
 * (sign = -1), GPUVoxelGrid.cpp:41-83 / 95-136, the (space, rgb, depth,
 * extr, withColor) overloads.  depth: device [H][pitch] uint16 mm; rgb:
 * device [H][W][3] uint8, or NULL to update sdf and weight only (the
 * reference's per-call withColor, GPUVoxelGrid.h:99-105); a grid created
 * without colour storage refuses a non-NULL rgb with TSDF_BAD_ARG. */
tsdf_status tsdf_integrate(tsdf_grid *g, const uint16_t *depth, int32_t pitch,
                           const uint8_t *rgb, const float E_cam2world[16],
                           int32_t sign, void *stream);

/* Ordered batch of integrate (+1) / deintegrate (-1) operations applied in
 * one pass over the affected bricks; bit-identical to calling
 * tsdf_integrate for op 0, 1, ..., n_ops-1 in turn.  Replaces the serial
 * loop of integrate.cpp:108-125 and the unused reintegratePP
 * (postprocessing.cpp:31-113).  depths/rgbs: HOST arrays of n_ops DEVICE
 * pointers (rgbs may be NULL, or hold NULL for an op without colour; a grid
 * without colour storage refuses a non-NULL rgbs); E: n_ops*16 host floats;
 * signs: n_ops. */
tsdf_status tsdf_integrate_batch(tsdf_grid *g, int32_t n_ops,
                                 const uint16_t *const *depths, int32_t pitch,
                                 const uint8_t *const *rgbs, const float *E,
                                 const int32_t *signs, void *stream);

/* Reintegration after a pose-graph update: for keyframe k = 0..n-1,
 * deintegrate with E_old[k] then integrate with E_new[k] (the order of
 * integrate.cpp:108-125), as one batch. */
tsdf_status tsdf_reintegrate_batch(tsdf_grid *g, int32_t n, const uint16_t *const *depths,
                                   int32_t pitch, const uint8_t *const *rgbs,
                                   const float *E_old, const float *E_new, void *stream);

/* Allocates the device buffers of a batch of up to n_ops ops (per-slot op
 * masks and one point map per op of a pass, at most 256 maps), so that later
 * batches of at most n_ops ops allocate nothing, and (with the tuning field
 * batch_graph) captures the graph of a batch of exactly n_ops under the
 * current tuning.  Without it the first batch allocates.  No reference
 * counterpart (the reference has no batch). */
tsdf_status tsdf_reserve_batch(tsdf_grid *g, int32_t n_ops);

/* GPUVoxelGrid::raycast(space, extr), GPUVoxelGrid.cpp:139-147 with
 * raycastingTag :561-715.  points: device [H][W][3] float, world space,
 * (0,0,0) on a miss (rayPcd).  colors: device [H][W][3] float in [0,1] or
 * NULL (rayColorPcd).  depth: device [H][W] float camera-frame z of the
 * point or NULL (extension).  ub_flags: device [H][W] uint8 or NULL. */
tsdf_status tsdf_raycast(tsdf_grid *g, const float E_cam2world[16], float *points,
                         float *colors, float *depth, uint8_t *ub_flags, void *stream);

/* ---- differentiable render (no reference counterpart) -----------------
 * The reference raycast (GPUVoxelGrid.cpp:561-715) is not differentiable;
 * these entry points render depth from the trilinear interpolation of the
 * sdf (first zero crossing along the ray, refined by Newton's method) and
 * give its exact derivative with respect to the pose and the voxel values
 * (implicit function theorem at the root).  They replace the point-to-plane
 * linearisation of SDF-to-pose tracking.  n_views <= 256 views share one
 * camera; E: n_views*16 host floats (camera->world, row-major); all image
 * buffers are DEVICE [n_views][h][w]; depth is camera-frame z in metres,
 * 0 where no surface is found.  sdf: optional DEVICE [n_bricks][512] float
 * values that replace the stored sdf (brick order of tsdf_export_bricks),
 * NULL = the stored values (an override needs a grid of at most 2^23 - 1
 * bricks); weights always come from the map. */

/* normals: DEVICE [n][h][w][3] unit gradient of the field, or NULL. */
tsdf_status tsdf_render(tsdf_grid *g, int32_t n_views, const float *E, const tsdf_view *cam,
                        const float *sdf, float *depth, float *normals, void *stream);

/* Adjoint of tsdf_render for a rendered `depth` and upstream gradient
 * grad_depth = dL/dD: grad_pose DEVICE [n][6] (overwritten) = dL/dxi for the
 * left perturbation E' = exp(xi^) E, xi = (translation, rotation);
 * grad_sdf DEVICE [n_bricks][512] (accumulated) = dL/dV, or NULL.  Sums over
 * pixels are taken in a fixed order: a repeated call gives the same bits.
 * The sdf override and grad_sdf address at most 2^23 - 1 bricks, and
 * grad_sdf at most 2^31 - 1 records (8 per pixel of all views):
 * TSDF_BAD_ARG beyond.  Grows grid-owned scratch (fp32 partials per
 * 16x16 pixels; with grad_sdf, 128 B of records per pixel) on first need. */
tsdf_status tsdf_render_backward(tsdf_grid *g, int32_t n_views, const float *E, const tsdf_view *cam,
                                 const float *sdf, const float *depth, const float *grad_depth,
                                 float *grad_pose, float *grad_sdf, void *stream);

/* Gauss-Newton normal equations of depth tracking against depth_obs
 * (DEVICE [n][h][w] metres, 0 = none): per view, over pixels where the
 * render hits, the observation is valid and |D - D_obs| < max_residual,
 * J = dD/dxi and r = D - D_obs; normal_eq DEVICE [n][29] doubles
 * (overwritten) = J^T J upper triangle row by row (21), J^T r (6), sum r^2,
 * pixel count.  depth: DEVICE [n][h][w] rendered depth, or NULL. */
tsdf_status tsdf_track_normal_eq(tsdf_grid *g, int32_t n_views, const float *E, const tsdf_view *cam,
                                 const float *depth_obs, float max_residual, double *normal_eq,
                                 float *depth, void *stream);

/* Host, fp64: solves J^T J delta = -J^T r (Cholesky of the 6x6 system given
 * as its upper triangle, 21 values as above) and composes
 * T_out = exp(delta^) T_in (T row-major camera->world).  delta: 6 or NULL.
 * TSDF_BAD_ARG when the system is not positive definite. */
tsdf_status tsdf_se3_gauss_newton_step(const double JtJ[21], const double Jtr[6], const double T_in[16],
                                       double T_out[16], double delta[6]);

/* Defaults, current values and replacement of the launch shapes. */
void tsdf_default_tuning(tsdf_tuning *t);
tsdf_status tsdf_get_tuning(const tsdf_grid *g, tsdf_tuning *t);
tsdf_status tsdf_set_tuning(tsdf_grid *g, const tsdf_tuning *t);

/* Diagnostic raycast from E that counts the distinct voxels read and the
 * distinct block keys probed (hits and misses) by the march of
 * tsdf_raycast: the compulsory-traffic terms of the raycast's roofline.
 * Synchronous; writes no image output. */
tsdf_status tsdf_raycast_footprint(tsdf_grid *g, const float E_cam2world[16],
                                   uint64_t *distinct_voxels, uint64_t *distinct_keys,
                                   void *stream);

/* Per-kernel timing: while enabled, every launch is bracketed by a pair of
 * HIP events on its stream.  tsdf_profile_read waits for the recorded events
 * and returns the summed device time per kernel kind (and clears the sums
 * when reset != 0). */
tsdf_status tsdf_profile(tsdf_grid *g, int32_t enable);
tsdf_status tsdf_profile_read(tsdf_grid *g, tsdf_kernel_times *out, int32_t reset);

/* Waits for `stream`, then reports TSDF_MAP_FULL if any allocation was
 * refused since the last reset. */
tsdf_status tsdf_sync(tsdf_grid *g, void *stream);

/* Waits for `stream` and reads the counters (voxelMap2.size(),
 * maxVoxelBlocks and the per-call counters). */
tsdf_status tsdf_stats(tsdf_grid *g, tsdf_stats_out *out, void *stream);

/* Copies the allocated bricks into caller-owned DEVICE buffers, converted to
 * fp32: coords [cap][3] int32 block indices, sdf/w [cap][512] (voxel
 * x + 8y + 64z, GPUVoxelGrid.h:28-31), rgb [cap][512][3] or NULL.  Writes
 * the brick count to *n_out (host) and copies min(n, cap) bricks.  The
 * reference equivalent is the deep_copy of voxelMap2 in marchingCubesHost,
 * GPUVoxelGrid.cpp:204-245. */
tsdf_status tsdf_export_bricks(tsdf_grid *g, int32_t *coords, float *sdf, float *w,
                               float *rgb, uint64_t cap, uint64_t *n_out, void *stream);

/* GPUVoxelGrid::cleanMemory(space), GPUVoxelGrid.cpp:185-196 with cleanTag
 * :785-805: erases every brick whose 512 weights are all zero (the surviving
 * bricks are compacted and the table rebuilt).  Synchronous; *n_removed (or
 * NULL) receives the number erased. */
tsdf_status tsdf_clean(tsdf_grid *g, uint64_t *n_removed, void *stream);

/* Triangle mesh of the zero level set (stands in for marchingCubesHost,
 * GPUVoxelGrid.cpp:200-251, which hands the bricks to Open3D's
 * ExtractTriangleMesh): marching tetrahedra over every cell of 8 observed
 * voxel centres, 6 tetrahedra per cell split around the cell diagonal, each
 * triangle facing the free-space side.  triangles: DEVICE [cap][3][3] float
 * world coordinates (a triangle soup); colors: DEVICE [cap][3][3] vertex
 * colours interpolated along the same edges, or NULL (a grid without colour
 * accepts only NULL); vertex_keys: DEVICE [cap][3] or NULL, a key per vertex
 * naming its edge (inside voxel's global index, biased by 2^18 per axis, in
 * bits 43-61 / 24-42 / 5-23, and the offset to the outside voxel,
 * (dx+1) + 3 (dy+1) + 9 (dz+1), in bits 0-4): equal keys are equal vertices,
 * which welds the soup into an indexed mesh (maps beyond +-32767 blocks per
 * axis are refused with keys); *n_triangles (host) receives the total, which
 * may exceed cap (then only cap are written).  Synchronous. */
tsdf_status tsdf_extract_mesh(tsdf_grid *g, float *triangles, float *colors, uint64_t *vertex_keys, uint64_t cap,
                              uint64_t *n_triangles, void *stream);

/* Marching cubes: the mesh of marchingCubesHost, GPUVoxelGrid.cpp:200-251
 * (Open3D 0.12 ScalableTSDFVolume::ExtractTriangleMesh over the bricks, then
 * ComputeVertexNormals), as an indexed mesh.  A cell of 8 voxels is meshed
 * when no corner has w == 0, by the Lorensen-Cline case tables; one vertex
 * per lattice edge (global voxel G, axis a), at vl/2 + vl G plus
 * f0 vl / (f0 + f1) along a (doubles, f = |sdf| at the edge's ends), with
 * colour (f1 c0 + f0 c1) / (f0 + f1) of c = rgb / 255; triangles face the
 * free-space side; normals are the normalised sums of the unnormalised face
 * normals (zero where the sum is zero).  DEVICE outputs: vertices
 * [vertex_cap][3] double; vertex_colors [vertex_cap][3] double or NULL (a
 * grid without colour accepts only NULL); vertex_normals [vertex_cap][3]
 * double or NULL; triangles [triangle_cap][3] int32 vertex indices;
 * vertex_keys [vertex_cap] or NULL, the edge of each vertex (G biased by
 * 2^18 per axis in bits 43-61 / 24-42 / 5-23, a in bits 0-4; maps beyond
 * +-32767 blocks per axis are refused).  *n_vertices and *n_triangles (host)
 * receive the mesh's sizes; nothing is written unless both fit their
 * capacities.  Vertex order is not fixed from call to call; the vertex set
 * and the triangle multiset are.  Synchronous. */
tsdf_status tsdf_marching_cubes(tsdf_grid *g, double *vertices, double *vertex_colors, double *vertex_normals,
                                int32_t *triangles, uint64_t *vertex_keys, uint64_t vertex_cap,
                                uint64_t triangle_cap, uint64_t *n_vertices, uint64_t *n_triangles, void *stream);

/* ---- point maps, normal maps and projective ICP (kokkosICP.cpp) ---------- */

typedef enum {
    TSDF_ICP_POINT_TO_POINT = 0, /* ICPTypes::PointToPoint, kokkosICP.h:4-8 */
    TSDF_ICP_POINT_TO_PLANE = 1  /* ICPTypes::PointToPlane                  */
} tsdf_icp_type;

typedef struct {
    int32_t type;         /* tsdf_icp_type                                   */
    int32_t n_iterations; /* n_it = 10, kokkosICP.cpp:49                     */
    int32_t n_search;     /* window half-width, n_search = 5, :48 (<= 64)    */
    float max_dist;       /* metres, max_dist = 0.075, :50                   */
} tsdf_icp_params;

/* Number of sums tsdf_icp_moments returns. */
#define TSDF_ICP_MOMENTS 18

void tsdf_icp_default_params(tsdf_icp_params *p);

/* CreatePcdFromDepthView, imageutil.cpp:89-111: points DEVICE [H][W][3] float
 * = ((j - cx) / fx d, (i - (H - cy)) / fy d, d), d = depth / 1000 (depth
 * DEVICE [H][W] uint16 mm, row-major).  Asynchronous on the stream. */
tsdf_status tsdf_depth_to_points(const uint16_t *depth, const tsdf_view *cam, float *points, void *stream);

/* CalculateNormalsFromPCDView, imageutil.cpp:116-157: normals DEVICE
 * [H][W][3] float; (0, 0, 1) on the border and where the pixel or a
 * 4-neighbour has z == 0, else the normalised sum of the four neighbour cross
 * products.  Asynchronous. */
tsdf_status tsdf_points_to_normals(const float *points, int32_t height, int32_t width, float *normals,
                                   void *stream);

/* The "Closest Point Kernel" of kokkosICP, kokkosICP.cpp:69-115: for every
 * pixel of x (z != 0) the closest point of y in the 2n x 2n window around the
 * projection of Txy x (row-major float Txy, the float cast the reference
 * uses), or its projection onto the tangent plane of y (point-to-plane, with
 * y_normals); x_nearest DEVICE [H][W][3], (0, 0, 0) where none qualifies.
 * Asynchronous. */
tsdf_status tsdf_icp_associate(const float *x_points, const float *y_points, const float *y_normals,
                               const float Txy[16], const tsdf_view *cam, const tsdf_icp_params *p,
                               float *x_nearest, void *stream);

/* The sums of one ICP iteration (kokkosICP.cpp:117-171), in double:
 * [0] points of x, [1] correspondences, [2..4] sum x, [5..7] sum x_nearest,
 * [8..16] sum x x_nearest^T (row-major), [17] sum |x_nearest - Txy x|^2 with
 * the double Txy.  moments: HOST [18].  Synchronous. */
tsdf_status tsdf_icp_moments(const float *x_points, const float *x_nearest, const double Txy[16],
                             int32_t height, int32_t width, double *moments, void *stream);

/* The sums of tsdf_icp_moments, bit-identical to them, for callers that
 * cannot allocate device memory (tight tracking loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It allocates nothing: its
 * partial sums live in a fixed 73 872-byte static device array (one per
 * device, held until the process ends), so it takes any image of at most
 * 2^29 - 1 pixels and refuses larger ones with TSDF_BAD_ARG.  Calls on one
 * device share that array: overlapping calls would race on it, so the
 * library serialises them with a per-device lock held until the call's
 * stream has drained.  Synchronous. */
tsdf_status tsdf_icp_moments_noalloc(const float *x_points, const float *x_nearest, const double Txy[16],
                                     int32_t height, int32_t width, double *moments, void *stream);

/* kokkosICP(space, d1, d2, tx, ty, type, &rsme), kokkosICP.cpp:23-191:
 * aligns the points of depth1 (pose tx, camera->world) to those of depth2
 * (pose ty), starting from Txy = ty^-1 tx; each iteration associates, takes
 * the centroids and the rotation of the SVD of the centred cross-covariance,
 * and the transform of the iteration with the lowest RMSE is returned in
 * T_out (host, row-major, x -> y); *rmse_out that RMSE (1e6 when no iteration
 * found a correspondence) and *fitness_out the best correspondences / points
 * ratio (either may be NULL).  Synchronous. */
tsdf_status tsdf_icp(const uint16_t *depth1, const uint16_t *depth2, const double tx[16], const double ty[16],
                     const tsdf_view *cam, const tsdf_icp_params *p, double T_out[16], float *rmse_out,
                     float *fitness_out, void *stream);

/* Frees the workspace tsdf_icp keeps on the current device between calls
 * (its point maps grow to the largest frame seen; no reference counterpart).
 * The next tsdf_icp on the device allocates it again. */
tsdf_status tsdf_icp_release(void);

/* ---- reintegration scheduler helpers (host only, no device work) ------- */

/* MattoDof, matrixutil.cpp:38-64: (rx, ry, rz, tx, ty, tz) of a row-major
 * camera->world T with R = Rz(rz) Ry(ry) Rx(rx). */
tsdf_status tsdf_mat_to_dof(const double T[16], double dof[6]);

/* getT, matrixutil.h:20-63: trans(t) Rz Ry Rx, evaluated left to right. */
tsdf_status tsdf_dof_to_mat(const double dof[6], double T[16]);

/* diff, integrate.cpp:211-216: |[world - integrated](0..2) , 2 [world -
 * integrated](3..5)| -- rotation angles, then twice the translation. */
double tsdf_reint_diff(const double world_dof[6], const double integrated_dof[6]);

/* updateReintegrationBuffer, integrate.cpp:223-234: order [n] receives the
 * list indices of the integrated frames (world_dof / integrated_dof: [n][6],
 * in list order) sorted by diff, largest first, equal ones in list order (the
 * list's new order); the first *n_selected of them, up to the first diff
 * below treint, form the reintegration buffer. */
tsdf_status tsdf_reint_select(int32_t n, const double *world_dof, const double *integrated_dof, double treint,
                              int32_t *order, int32_t *n_selected);

/* Host helper: the fp32 pose data every kernel uses -- E^-1 of a rigid pose
 * evaluated in double and rounded once (stand-in for Eigen's float inverse,
 * GPUVoxelGrid.cpp:55) and Ez = normalize(third column) (:52-53). */
tsdf_status tsdf_pose_prepare(const float E_cam2world[16], float E_inv[16], float Ez[3]);

const char *tsdf_status_string(tsdf_status s);

/* Text of the last error on this thread ("" if none). */
const char *tsdf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
