/*
 * tsdf_oracle.h -- CPU oracle for the voxel-hashed TSDF path
 * (block allocation -> projective integrate / deintegrate -> raycast).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and nowhere else.  The product (superscanner_amd, libtsdf_hip.so) never
 * links, loads or calls it.
 *
 * It restates, in plain C with binary32 arithmetic in source order (built with
 * -ffp-contract=off), the algorithm of the reference files
 *   src/VoxelGrid/GPUVoxelGrid.h     :11-38, 124-307
 *   src/VoxelGrid/GPUVoxelGrid.cpp   :41-147, 253-266, 345-722
 *   src/VoxelGrid/kokkosFunctions.h  :90-268, 394-483, 737-740
 * Every function in tsdf_oracle.c cites the line it follows.
 */
#ifndef TSDF_ORACLE_H
#define TSDF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;

typedef struct {
    int64_t n_touched; /* distinct bricks in this call's touched set   */
    int64_t n_new;     /* bricks inserted by this call                  */
    int64_t n_upd;     /* voxels whose (sdf, w) this call changed       */
    int64_t n_valid;   /* pixels with camera z > rmin                   */
} orc_counts;

typedef struct {
    int64_t n_hit;     /* pixels with a zero crossing                   */
    int64_t n_ub;      /* pixels whose march hit reference UB (masked)  */
    int64_t samples;   /* voxel reads over all pixels                   */
    int64_t probes;    /* hash lookups over all pixels                  */
    int64_t u_vox;     /* distinct voxels read (stats mode only)        */
    int64_t u_blk;     /* distinct keys probed, hits + misses (stats)   */
} orc_ray_counts;

/* UB flags written per pixel by orc_raycast (0 = well defined). */
#define ORC_UB_VOXEL_INDEX 1  /* getVinVBIndeces gave -1 or 8 (GPUVoxelGrid.h:192-221) */
#define ORC_UB_MISSING_OUT 2  /* v_out read from a block that is not in the map (.cpp:669-672) */

/* --- scalar kernels (exported so tests can pin them with KATs) --- */
uint64_t orc_prime_hash(int x, int y, int z);
int      orc_f2i(float f);
int      orc_vb_index(float a, float vbl);
void     orc_vin_vb(const float p[3], const int vb[3], float vbl, float vl, int out[3]);
void     orc_pose_prepare(const float E[16], float Einv[16], float Ez[3]);
int      orc_pose_is_rigid(const float E[16]);

/* --- grid --- */
orc_grid *orc_create(float voxel_length, float truncation, int height, int width,
                     int64_t max_bricks, int with_color, float max_weight,
                     float min_depth, float max_ray_length);
void     orc_destroy(orc_grid *g);
void     orc_set_intrinsics(orc_grid *g, float fx, float fy, float cx, float cy);
void     orc_set_threads(int n);
/* Test-only: plants a deliberate mistake (0 = none, 1 sdf sign, 2 pose not
 * transposed, 3 floor block index, 4 textbook ci, 5 flipped interpolation,
 * 6 deintegrate keeps sdf at w == 1, 7 render adjoint: rotation part with
 * p x grad phi reversed, 8 render adjoint: voxel gradient without the 1/s
 * factor). */
void     orc_set_mutation(int m);
int      orc_mutation(void);

/* sign = +1 integrate, -1 deintegrate.  Returns 0, or 3 when the brick pool
 * overflowed (then the grid is no longer a valid reference state). */
int      orc_integrate(orc_grid *g, const uint16_t *depth, int pitch,
                       const uint8_t *rgb, const float E[16], int sign,
                       orc_counts *out);

/* points/colors: H*W*3 floats (colors may be NULL), depth: H*W floats or NULL,
 * ub: H*W bytes or NULL.  stats != 0 additionally fills u_vox / u_blk
 * (single threaded). */
void     orc_raycast(orc_grid *g, const float E[16], float *points, float *colors,
                     float *depth, uint8_t *ub, int stats, orc_ray_counts *out);

int64_t  orc_n_bricks(const orc_grid *g);
/* cleanMemory (GPUVoxelGrid.cpp:185-196): erases the bricks whose weights are
 * all zero; returns how many were erased. */
int64_t  orc_clean(orc_grid *g);
void     orc_import_rgb(orc_grid *g, const float *rgb);
/* coords: n*3 int32, keys: n int64 (sign-extended reference hash),
 * sdf/w: n*512 floats, rgb: n*512*3 floats or NULL.  Brick order = insertion
 * order; voxel order x + 8y + 64z (GPUVoxelGrid.h:28-31). */
void     orc_export(const orc_grid *g, int32_t *coords, int64_t *keys,
                    float *sdf, float *w, float *rgb);

/* --- differentiable trilinear render (render_oracle.c), fp64 ---------
 * No reference counterpart (the reference raycast is not differentiable);
 * the definition of the field, march, root and adjoint is in
 * render_oracle.c.  A field is built from exported bricks (coords n*3,
 * sdf / w n*512, brick order of the export). */
typedef struct orc_field orc_field;
orc_field *orc_field_create(int64_t n, const int32_t *coords, const float *sdf, const float *w,
                            float vl, float rmin, float dmax);
void     orc_field_destroy(orc_field *f);
void     orc_field_set_sdf(orc_field *f, const float *sdf);
/* depth: h*w (0 = no surface), normal: h*w*3 or NULL.  E: camera->world,
 * row-major, double. */
void     orc_render(const orc_field *f, const double E[16], int h, int w, double fx, double fy,
                    double cx, double cy, double *depth, double *normal);
/* For the rendered depth: jac (h*w*6 or NULL) = dD/dxi per pixel (left
 * perturbation xi = (v, w) of E), grad_pose (6 or NULL) = sum of
 * grad_depth * jac, grad_sdf (n*512 or NULL) += grad_depth * dD/dV. */
void     orc_render_grad(const orc_field *f, const double E[16], int h, int w, double fx, double fy,
                         double cx, double cy, const double *depth, const double *grad_depth,
                         double *jac, double *grad_pose, double *grad_sdf);

/* --- fp64 integrate on a crop (fp64_oracle.c) ------------------------
 * The integrate of orc_integrate evaluated in double, keeping only blocks
 * whose box meets [crop_min, crop_max]; measures the error of the fp32 /
 * fp16 GPU grids.  E: camera->world, row-major, double. */
typedef struct orc64_grid orc64_grid;
orc64_grid *orc64_create(double vl, double trunc, int H, int W, double fx, double fy, double cx, double cy,
                         const double crop_min[3], const double crop_max[3], double max_weight,
                         double min_depth);
void     orc64_destroy(orc64_grid *g);
void     orc64_integrate(orc64_grid *g, const uint16_t *depth, int pitch, const double E[16], int sign);
int64_t  orc64_n_bricks(const orc64_grid *g);
/* coords n*3 int32, sdf / w n*512 double (voxel x + 8y + 64z). */
void     orc64_export(const orc64_grid *g, int32_t *coords, double *sdf, double *w);

/* icp_oracle.c: point maps, normal maps, projective ICP (imageutil.cpp,
 * kokkosICP.cpp) */
void orc_depth_to_points(const uint16_t *depth, int H, int W, float fx, float fy, float cx, float cy, float *out);
void orc_normals(const float *pcd, int H, int W, float *out);
void orc_icp_associate(const float *x, const float *y, const float *yn, const float *T, int H, int W, float fx,
                       float fy, float cx, float cy, int type, int n_search, float max_dist, float *out);
void orc_icp_sums(const float *x, const float *xn, const double *T, int H, int W, double *out);

#ifdef __cplusplus
}
#endif
#endif
