/*
 * render_oracle.c -- fp64 CPU reference of the differentiable trilinear
 * render of a TSDF and of its adjoint.
 *
 * TEST INFRASTRUCTURE ONLY (like tsdf_oracle.c): loaded by tests/ and by the
 * cpu_baseline leg of bench.py, never by the product.
 *
 * No reference counterpart: the reference raycast (GPUVoxelGrid.cpp:561-715)
 * samples voxel centres and is not differentiable.  This is the extension's
 * own definition, restated here in double precision; the HIP kernels in
 * superscanner_amd/csrc/tsdf_render.hip implement the same steps in fp32.
 *
 * Field.  Voxel k (global index, k = 8 * block + v) has its centre at
 * (k + 0.5) * vl.  phi(p) is the trilinear interpolation of the sdf values of
 * the 8 voxel centres around p; it is defined ("valid") only when all 8
 * voxels exist and carry a weight > 0.  Bricks are looked up by the
 * reference key (GPUVoxelGrid.h:124-128), as every other lookup of the map.
 *
 * Camera.  Pixel (i, j) of an h x w view with intrinsics (fx, fy, cx, cy)
 * looks along r = ((j - cx) / fx, (i - (h - cy)) / fy, 1) in camera
 * coordinates (the reference convention ci = H - cy, cj = cx,
 * GPUVoxelGrid.cpp:253-266); the world ray is o + lam * R r, so lam is the
 * camera-frame depth.
 *
 * March.  From lam = rmin: at a valid sample step by
 * max(0.5 |phi|, 0.5 vl) / |R r|; at an invalid sample inside an allocated
 * block step by 0.5 vl / |R r|; in a block that is not in the map (where phi
 * is invalid everywhere) jump to the block's exit.  The first pair of
 * consecutive valid samples with phi_prev > 0 >= phi brackets the surface.
 *
 * Root.  Safeguarded Newton on g(lam) = phi(o + lam R r) inside the bracket
 * (bisection when a step leaves it) until the step is below tol.  A hit
 * needs dphi/dlam < -kGraze |grad phi| |R r|.
 *
 * Adjoint (implicit function theorem at g(lam*) = 0, s = grad phi . R r):
 *   pose, left perturbation T' = exp(xi^) T, xi = (v, w):
 *     dD/dv = -grad phi / s,  dD/dw = -(p* x grad phi) / s
 *   voxel c of the 8 around p*, trilinear weight a_c:
 *     dD/dV_c = -a_c / s
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "tsdf_oracle.h"

#define VPE 8
#define VPB 512

static const double kStepFrac = 0.5;   /* fraction of |phi| per march step   */
static const double kMinStepVox = 0.5; /* minimum step, in voxels            */
static const double kGraze = 0.05;     /* minimum |cos| between ray and grad */
static const int kMaxIter = 48;

struct orc_field {
    int64_t n;
    float vl, rmin, dmax;
    int32_t *coords;
    float *sdf, *w;
    uint64_t hcap;
    uint32_t *hkey;
    int64_t *hval;
};

static inline uint32_t mix32(uint32_t h)
{
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

static inline uint32_t key_of(int64_t x, int64_t y, int64_t z)
{
    return (uint32_t)orc_prime_hash((int)x, (int)y, (int)z);
}

static int64_t fh_find(const orc_field *f, uint32_t k)
{
    uint64_t m = f->hcap - 1, b = mix32(k) & m;
    while (f->hval[b] >= 0) {
        if (f->hkey[b] == k)
            return f->hval[b];
        b = (b + 1) & m;
    }
    return -1;
}

orc_field *orc_field_create(int64_t n, const int32_t *coords, const float *sdf, const float *w,
                            float vl, float rmin, float dmax)
{
    orc_field *f = (orc_field *)calloc(1, sizeof(orc_field));
    f->n = n;
    f->vl = vl;
    f->rmin = rmin;
    f->dmax = dmax;
    f->coords = (int32_t *)malloc((size_t)(n ? n : 1) * 3 * sizeof(int32_t));
    f->sdf = (float *)malloc((size_t)(n ? n : 1) * VPB * sizeof(float));
    f->w = (float *)malloc((size_t)(n ? n : 1) * VPB * sizeof(float));
    memcpy(f->coords, coords, (size_t)n * 3 * sizeof(int32_t));
    memcpy(f->sdf, sdf, (size_t)n * VPB * sizeof(float));
    memcpy(f->w, w, (size_t)n * VPB * sizeof(float));
    f->hcap = 1024;
    while (f->hcap < (uint64_t)(4 * n))
        f->hcap <<= 1;
    f->hkey = (uint32_t *)malloc(f->hcap * sizeof(uint32_t));
    f->hval = (int64_t *)malloc(f->hcap * sizeof(int64_t));
    for (uint64_t i = 0; i < f->hcap; i++)
        f->hval[i] = -1;
    for (int64_t b = 0; b < n; b++) {
        uint32_t k = key_of(coords[3 * b], coords[3 * b + 1], coords[3 * b + 2]);
        uint64_t m = f->hcap - 1, s = mix32(k) & m;
        while (f->hval[s] >= 0 && f->hkey[s] != k)
            s = (s + 1) & m;
        if (f->hval[s] < 0) { /* first brick with this key wins, as in the map */
            f->hkey[s] = k;
            f->hval[s] = b;
        }
    }
    return f;
}

void orc_field_destroy(orc_field *f)
{
    if (!f)
        return;
    free(f->coords); free(f->sdf); free(f->w); free(f->hkey); free(f->hval);
    free(f);
}

void orc_field_set_sdf(orc_field *f, const float *sdf)
{
    memcpy(f->sdf, sdf, (size_t)f->n * VPB * sizeof(float));
}

static inline int64_t floor_div8(int64_t k) { return k >= 0 ? k / 8 : -((-k + 7) / 8); }

/* Trilinear value and gradient at p; returns 0 when the field is undefined.
 * idx / wt (optional) receive the 8 voxel indices (brick * 512 + v) and
 * trilinear weights, corner c = (c & 1, c >> 1 & 1, c >> 2). */
static int field_sample(const orc_field *f, const double p[3], double *phi, double grad[3], int64_t idx[8],
                        double wt[8])
{
    const double vl = f->vl;
    double fr[3];
    int64_t k0[3];
    for (int a = 0; a < 3; a++) {
        double u = p[a] / vl - 0.5;
        double fl = floor(u);
        k0[a] = (int64_t)fl;
        fr[a] = u - fl;
    }
    double V[8];
    int64_t cache_b = -2, cb[3] = {0, 0, 0};
    for (int c = 0; c < 8; c++) {
        int64_t k[3] = {k0[0] + (c & 1), k0[1] + ((c >> 1) & 1), k0[2] + ((c >> 2) & 1)};
        int64_t bk[3] = {floor_div8(k[0]), floor_div8(k[1]), floor_div8(k[2])};
        int64_t b;
        if (cache_b != -2 && bk[0] == cb[0] && bk[1] == cb[1] && bk[2] == cb[2]) {
            b = cache_b;
        } else {
            b = fh_find(f, key_of(bk[0], bk[1], bk[2]));
            cache_b = b;
            cb[0] = bk[0]; cb[1] = bk[1]; cb[2] = bk[2];
        }
        if (b < 0)
            return 0;
        int64_t v = (k[0] - 8 * bk[0]) + 8 * (k[1] - 8 * bk[1]) + 64 * (k[2] - 8 * bk[2]);
        int64_t id = b * VPB + v;
        if (!(f->w[id] > 0.0f))
            return 0;
        V[c] = (double)f->sdf[id];
        if (idx)
            idx[c] = id;
    }
    double s = 0, g0 = 0, g1 = 0, g2 = 0;
    for (int c = 0; c < 8; c++) {
        double ax = (c & 1) ? fr[0] : 1.0 - fr[0];
        double ay = ((c >> 1) & 1) ? fr[1] : 1.0 - fr[1];
        double az = ((c >> 2) & 1) ? fr[2] : 1.0 - fr[2];
        double sx = (c & 1) ? 1.0 : -1.0, sy = ((c >> 1) & 1) ? 1.0 : -1.0, sz = ((c >> 2) & 1) ? 1.0 : -1.0;
        s += ax * ay * az * V[c];
        g0 += sx * ay * az * V[c];
        g1 += ax * sy * az * V[c];
        g2 += ax * ay * sz * V[c];
        if (wt)
            wt[c] = ax * ay * az;
    }
    *phi = s;
    if (grad) {
        grad[0] = g0 / vl;
        grad[1] = g1 / vl;
        grad[2] = g2 / vl;
    }
    return 1;
}

static int block_exists(const orc_field *f, const double p[3], double vbl, int64_t bk[3])
{
    for (int a = 0; a < 3; a++)
        bk[a] = (int64_t)floor(p[a] / vbl);
    return fh_find(f, key_of(bk[0], bk[1], bk[2])) >= 0;
}

typedef struct {
    double o[3], d[3], dn;
} ray_t;

static ray_t make_ray(const double E[16], int h, double fx, double fy, double cx, double cy, int i, int j)
{
    ray_t r;
    const double ci = (double)h - cy, cj = cx;
    const double rc[3] = {((double)j - cj) / fx, ((double)i - ci) / fy, 1.0};
    for (int a = 0; a < 3; a++) {
        r.o[a] = E[4 * a + 3];
        r.d[a] = E[4 * a] * rc[0] + E[4 * a + 1] * rc[1] + E[4 * a + 2] * rc[2];
    }
    r.dn = sqrt(r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2]);
    return r;
}

static inline void at(const ray_t *r, double lam, double p[3])
{
    for (int a = 0; a < 3; a++)
        p[a] = r->o[a] + lam * r->d[a];
}

/* Depth of the first surface along the ray (0 = none). */
static double trace(const orc_field *f, const ray_t *r)
{
    const double vl = f->vl, vbl = 8.0 * vl;
    double lam = f->rmin, prev_lam = 0, prev_phi = 0;
    int prev_valid = 0;
    double p[3], phi = 0, grad[3];
    while (lam < f->dmax) {
        at(r, lam, p);
        int valid = field_sample(f, p, &phi, NULL, NULL, NULL);
        if (valid && prev_valid && prev_phi > 0 && phi <= 0) {
            /* safeguarded Newton inside [a, b] */
            double a = prev_lam, b = lam, fa = prev_phi, fb = phi;
            double x = (fb == 0) ? b : b - fb * (b - a) / (fb - fa);
            for (int it = 0; it < kMaxIter; it++) {
                at(r, x, p);
                double fx_;
                if (!field_sample(f, p, &fx_, grad, NULL, NULL))
                    return 0;
                if (fx_ == 0)
                    break;
                if (fx_ > 0)
                    a = x;
                else
                    b = x;
                double s = grad[0] * r->d[0] + grad[1] * r->d[1] + grad[2] * r->d[2];
                double nx = (s < 0) ? x - fx_ / s : 0.5 * (a + b);
                if (!(nx > a && nx < b))
                    nx = 0.5 * (a + b);
                double dx = fabs(nx - x);
                x = nx;
                if (dx <= 1e-13 * (1.0 + x))
                    break;
            }
            at(r, x, p);
            if (!field_sample(f, p, &phi, grad, NULL, NULL))
                return 0;
            double s = grad[0] * r->d[0] + grad[1] * r->d[1] + grad[2] * r->d[2];
            double gn = sqrt(grad[0] * grad[0] + grad[1] * grad[1] + grad[2] * grad[2]);
            if (!(s < -kGraze * gn * r->dn))
                return 0;
            return x;
        }
        double step;
        if (valid) {
            double m = kStepFrac * fabs(phi);
            step = (m > kMinStepVox * vl ? m : kMinStepVox * vl) / r->dn;
        } else {
            int64_t bk[3];
            if (block_exists(f, p, vbl, bk)) {
                step = kMinStepVox * vl / r->dn;
            } else {
                /* leave the empty block: exit parameter of its box */
                double te = INFINITY;
                for (int a2 = 0; a2 < 3; a2++) {
                    if (r->d[a2] > 0)
                        te = fmin(te, ((double)(bk[a2] + 1) * vbl - r->o[a2]) / r->d[a2]);
                    else if (r->d[a2] < 0)
                        te = fmin(te, ((double)bk[a2] * vbl - r->o[a2]) / r->d[a2]);
                }
                step = te - lam + 1e-3 * vl / r->dn;
                if (!(step > 1e-3 * vl / r->dn))
                    step = 1e-3 * vl / r->dn;
            }
        }
        prev_lam = lam;
        prev_phi = phi;
        prev_valid = valid;
        lam += step;
    }
    return 0;
}

void orc_render(const orc_field *f, const double E[16], int h, int w, double fx, double fy, double cx, double cy,
                double *depth, double *normal)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < h; i++) {
        for (int j = 0; j < w; j++) {
            ray_t r = make_ray(E, h, fx, fy, cx, cy, i, j);
            double lam = trace(f, &r);
            depth[(int64_t)i * w + j] = lam;
            if (normal) {
                double *nm = normal + 3 * ((int64_t)i * w + j);
                nm[0] = nm[1] = nm[2] = 0;
                if (lam > 0) {
                    double p[3], phi, g[3];
                    at(&r, lam, p);
                    if (field_sample(f, p, &phi, g, NULL, NULL)) {
                        double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                        nm[0] = g[0] / gn;
                        nm[1] = g[1] / gn;
                        nm[2] = g[2] / gn;
                    }
                }
            }
        }
    }
}

void orc_render_grad(const orc_field *f, const double E[16], int h, int w, double fx, double fy, double cx,
                     double cy, const double *depth, const double *grad_depth, double *jac, double *grad_pose,
                     double *grad_sdf)
{
    double gp[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < h; i++) {
        for (int j = 0; j < w; j++) {
            const int64_t u = (int64_t)i * w + j;
            double J[6] = {0, 0, 0, 0, 0, 0};
            if (depth[u] > 0) {
                ray_t r = make_ray(E, h, fx, fy, cx, cy, i, j);
                double p[3], phi, g[3], wt[8];
                int64_t idx[8];
                at(&r, depth[u], p);
                if (field_sample(f, p, &phi, g, idx, wt)) {
                    double s = g[0] * r.d[0] + g[1] * r.d[1] + g[2] * r.d[2];
                    double cxg[3] = {p[1] * g[2] - p[2] * g[1], p[2] * g[0] - p[0] * g[2],
                                     p[0] * g[1] - p[1] * g[0]};
                    if (orc_mutation() == 7)
                        for (int a = 0; a < 3; a++)
                            cxg[a] = -cxg[a];
                    for (int a = 0; a < 3; a++) {
                        J[a] = -g[a] / s;
                        J[3 + a] = -cxg[a] / s;
                    }
                    const double gd = grad_depth ? grad_depth[u] : 0.0;
                    for (int k = 0; k < 6; k++)
                        gp[k] += gd * J[k];
                    if (grad_sdf)
                        for (int c = 0; c < 8; c++)
                            grad_sdf[idx[c]] += gd * (orc_mutation() == 8 ? -wt[c] : -wt[c] / s);
                }
            }
            if (jac)
                for (int k = 0; k < 6; k++)
                    jac[6 * u + k] = J[k];
        }
    }
    if (grad_pose)
        for (int k = 0; k < 6; k++)
            grad_pose[k] = gp[k];
}
