/*
 * fp64_oracle.c -- the TSDF integrate of the reference evaluated in double
 * precision, restricted to the blocks that intersect a crop box.
 *
 * TEST INFRASTRUCTURE ONLY (like tsdf_oracle.c): loaded by tests/, never by
 * the product.  It measures the error of the fp32-arithmetic / fp16-storage
 * GPU grid against exact-as-possible arithmetic on the same frames.
 *
 * Same steps as orc_integrate (GPUVoxelGrid.cpp:471-550, cited there), with
 * every quantity in double: the pixel point getPoint (GPUVoxelGrid.h:229-232)
 * from depth / 1000, its world point E p, the block range of the +-trunc box
 * with the reference's index rule a >= 0 ? (int)(a / vbl) : (int)(a / vbl - 1)
 * (GPUVoxelGrid.h:162-190), voxel centres pzero + (k + 0.5) vl, the
 * projection with E^-1 (rigid inverse in double) and (int) truncation, the
 * sdf (E vpt - vm) . Ez and the running average with the project's weight
 * cap.  Only blocks whose box meets the crop are allocated and updated.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "tsdf_oracle.h"

#define VPB 512

struct orc64_grid {
    double vl, vbl, trunc, rmin, wmax;
    int H, W;
    double fx, fy, ci, cj;
    int64_t crop_lo[3], crop_hi[3]; /* block index range of the crop, inclusive */
    int64_t n, cap;
    int32_t *coord;
    double *sdf, *w;
    uint32_t *stamp;
    uint32_t epoch;
    uint64_t hcap;
    int64_t *hkey; /* packed coords, -1 = empty */
    int64_t *hval;
};

static inline uint64_t mix64(uint64_t h)
{
    h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull; h ^= h >> 33;
    return h;
}

static inline int64_t pack(int64_t x, int64_t y, int64_t z)
{
    return ((x + (1 << 20)) << 42) | ((y + (1 << 20)) << 21) | (z + (1 << 20));
}

static int64_t g64_find(const struct orc64_grid *g, int64_t k)
{
    uint64_t m = g->hcap - 1, b = mix64((uint64_t)k) & m;
    while (g->hkey[b] != -1) {
        if (g->hkey[b] == k)
            return g->hval[b];
        b = (b + 1) & m;
    }
    return -1;
}

static void g64_put(struct orc64_grid *g, int64_t k, int64_t v)
{
    uint64_t m = g->hcap - 1, b = mix64((uint64_t)k) & m;
    while (g->hkey[b] != -1)
        b = (b + 1) & m;
    g->hkey[b] = k;
    g->hval[b] = v;
}

static int64_t vbi(double a, double vbl)
{
    return a >= 0 ? (int64_t)(a / vbl) : (int64_t)(a / vbl - 1.0);
}

orc64_grid *orc64_create(double vl, double trunc, int H, int W, double fx, double fy, double cx, double cy,
                         const double crop_min[3], const double crop_max[3], double max_weight, double min_depth)
{
    struct orc64_grid *g = (struct orc64_grid *)calloc(1, sizeof(*g));
    g->vl = vl;
    g->vbl = 8.0 * vl;
    g->trunc = trunc;
    g->rmin = min_depth;
    g->wmax = max_weight;
    g->H = H;
    g->W = W;
    g->fx = fx;
    g->fy = fy;
    g->ci = (double)H - cy;
    g->cj = cx;
    for (int a = 0; a < 3; a++) {
        g->crop_lo[a] = (int64_t)floor(crop_min[a] / g->vbl);
        g->crop_hi[a] = (int64_t)floor(crop_max[a] / g->vbl);
    }
    g->hcap = 1 << 16;
    g->hkey = (int64_t *)malloc(g->hcap * sizeof(int64_t));
    g->hval = (int64_t *)malloc(g->hcap * sizeof(int64_t));
    for (uint64_t i = 0; i < g->hcap; i++)
        g->hkey[i] = -1;
    return g;
}

void orc64_destroy(orc64_grid *g)
{
    if (!g)
        return;
    free(g->coord); free(g->sdf); free(g->w); free(g->stamp); free(g->hkey); free(g->hval);
    free(g);
}

static int64_t g64_insert(struct orc64_grid *g, int64_t x, int64_t y, int64_t z)
{
    int64_t k = pack(x, y, z);
    int64_t b = g64_find(g, k);
    if (b >= 0)
        return b;
    if (g->n == g->cap) {
        g->cap = g->cap ? 2 * g->cap : 1024;
        g->coord = (int32_t *)realloc(g->coord, g->cap * 3 * sizeof(int32_t));
        g->sdf = (double *)realloc(g->sdf, g->cap * VPB * sizeof(double));
        g->w = (double *)realloc(g->w, g->cap * VPB * sizeof(double));
        g->stamp = (uint32_t *)realloc(g->stamp, g->cap * sizeof(uint32_t));
    }
    b = g->n++;
    g->coord[3 * b] = (int32_t)x;
    g->coord[3 * b + 1] = (int32_t)y;
    g->coord[3 * b + 2] = (int32_t)z;
    memset(g->sdf + b * VPB, 0, VPB * sizeof(double));
    memset(g->w + b * VPB, 0, VPB * sizeof(double));
    g->stamp[b] = 0;
    if ((uint64_t)g->n * 2 > g->hcap) {
        uint64_t oc = g->hcap;
        int64_t *ok = g->hkey, *ov = g->hval;
        g->hcap *= 2;
        g->hkey = (int64_t *)malloc(g->hcap * sizeof(int64_t));
        g->hval = (int64_t *)malloc(g->hcap * sizeof(int64_t));
        for (uint64_t i = 0; i < g->hcap; i++)
            g->hkey[i] = -1;
        for (uint64_t i = 0; i < oc; i++)
            if (ok[i] != -1)
                g64_put(g, ok[i], ov[i]);
        free(ok);
        free(ov);
    }
    g64_put(g, k, b);
    return b;
}

static inline void point(const struct orc64_grid *g, const uint16_t *depth, int pitch, int i, int j, double p[3])
{
    double d = (double)depth[(int64_t)i * pitch + j] / 1000.0;
    p[0] = ((double)j - g->cj) / g->fx * d;
    p[1] = ((double)i - g->ci) / g->fy * d;
    p[2] = d;
}

static inline void xf(const double *E, const double p[3], double o[3])
{
    for (int r = 0; r < 3; r++)
        o[r] = E[4 * r] * p[0] + E[4 * r + 1] * p[1] + E[4 * r + 2] * p[2] + E[4 * r + 3];
}

void orc64_integrate(orc64_grid *g, const uint16_t *depth, int pitch, const double E[16], int sign)
{
    double Ei[16], Ez[3];
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++)
            Ei[4 * r + c] = E[4 * c + r];
        Ei[4 * r + 3] = -(E[r] * E[3] + E[4 + r] * E[7] + E[8 + r] * E[11]);
    }
    double len = sqrt(E[2] * E[2] + E[6] * E[6] + E[10] * E[10]);
    Ez[0] = E[2] / len;
    Ez[1] = E[6] / len;
    Ez[2] = E[10] / len;
    g->epoch++;
    int64_t tcap = 1024, tn = 0;
    int64_t *touched = (int64_t *)malloc(tcap * sizeof(int64_t));
    for (int i = 0; i < g->H; i++)
        for (int j = 0; j < g->W; j++) {
            double p[3], pt[3];
            point(g, depth, pitch, i, j, p);
            if (!(p[2] > g->rmin))
                continue;
            xf(E, p, pt);
            int64_t lo[3], hi[3];
            int skip = 0;
            for (int a = 0; a < 3; a++) {
                lo[a] = vbi(pt[a] - g->trunc, g->vbl);
                hi[a] = vbi(pt[a] + g->trunc, g->vbl);
                if (lo[a] < g->crop_lo[a])
                    lo[a] = g->crop_lo[a];
                if (hi[a] > g->crop_hi[a])
                    hi[a] = g->crop_hi[a];
                if (lo[a] > hi[a])
                    skip = 1;
            }
            if (skip)
                continue;
            for (int64_t x = lo[0]; x <= hi[0]; x++)
                for (int64_t y = lo[1]; y <= hi[1]; y++)
                    for (int64_t z = lo[2]; z <= hi[2]; z++) {
                        int64_t b = g64_insert(g, x, y, z);
                        if (g->stamp[b] != g->epoch) {
                            g->stamp[b] = g->epoch;
                            if (tn == tcap) {
                                tcap *= 2;
                                touched = (int64_t *)realloc(touched, tcap * sizeof(int64_t));
                            }
                            touched[tn++] = b;
                        }
                    }
        }
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t t = 0; t < tn; t++) {
        int64_t b = touched[t];
        const int32_t *c = g->coord + 3 * b;
        for (int v = 0; v < VPB; v++) {
            int k[3] = {v & 7, (v >> 3) & 7, v >> 6};
            double vm[3], cc[3];
            for (int a = 0; a < 3; a++)
                vm[a] = (double)c[a] * g->vbl + ((double)k[a] + 0.5) * g->vl;
            xf(Ei, vm, cc);
            double row = cc[1] * g->fy / cc[2] + g->ci, col = cc[0] * g->fx / cc[2] + g->cj;
            if (!(row > -1.0 && row < (double)g->H && col > -1.0 && col < (double)g->W))
                continue;
            int ri = (int)row, cj = (int)col;
            double p[3], vp[3];
            point(g, depth, pitch, ri, cj, p);
            if (!(p[2] > g->rmin))
                continue;
            xf(E, p, vp);
            double sdf = (vp[0] - vm[0]) * Ez[0] + (vp[1] - vm[1]) * Ez[1] + (vp[2] - vm[2]) * Ez[2];
            if (!(fabs(sdf) < g->trunc))
                continue;
            double *s = g->sdf + b * VPB + v, *w = g->w + b * VPB + v;
            double ww = *w;
            if (sign > 0) {
                *s = (*s * ww + sdf) / (ww + 1.0);
                *w = (g->wmax > 0 && ww >= g->wmax) ? g->wmax : ww + 1.0;
            } else if (ww != 1.0) {
                *s = (*s * ww - sdf) / (ww - 1.0);
                *w = ww - 1.0;
            } else {
                *s = 0.0;
                *w = 0.0;
            }
        }
    }
    free(touched);
}

int64_t orc64_n_bricks(const orc64_grid *g) { return g->n; }

void orc64_export(const orc64_grid *g, int32_t *coords, double *sdf, double *w)
{
    memcpy(coords, g->coord, (size_t)g->n * 3 * sizeof(int32_t));
    memcpy(sdf, g->sdf, (size_t)g->n * VPB * sizeof(double));
    memcpy(w, g->w, (size_t)g->n * VPB * sizeof(double));
}
