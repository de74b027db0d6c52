/*
 * tsdf_oracle.c -- CPU oracle for the voxel-hashed TSDF path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  The product never calls it.
 *
 * Plain C, binary32 in source order, compiled with -ffp-contract=off so that
 * no a*b+c is fused.  The C++ promotions of the reference are reproduced
 * explicitly: `(k + 0.5) * vl` and getVoxelMidPoint are evaluated in double
 * and rounded to float; the raycast colour uses `0.5 *` in double.
 *
 * Parity pinning (the tests that hold each function, see tests/):
 *   orc_prime_hash        closed-form KATs incl. the wrap-around and the
 *                         (152,951,211) alias               test_oracle_kat.py
 *   orc_vb_index          KATs at block boundaries incl. the -2 quirk
 *   orc_vin_vb            KATs + brute-force agreement with floor on the
 *                         interior of a block
 *   orc_pose_prepare      closed form (R^T, -R^T t) on random rigid poses
 *   orc_integrate         closed form on a fronto-parallel plane (sdf = D - z,
 *                         w = #frames), brute-force block enumeration of the
 *                         allocated set, integrate∘deintegrate restores w,
 *                         planted-mistake checks       test_oracle_closed_form.py
 *   orc_raycast           closed-form plane / sphere hit distance within one
 *                         voxel, miss -> 0, UB pixels excluded
 * The reference itself cannot be built here (Kokkos-CUDA only, Kokkos / Eigen
 * / Open3D absent), so there is no oracle/_ref.
 */
#include "tsdf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VPE 8    /* voxelPerEdge, GPUVoxelGrid.h:6 */
/* Deliberate-mistake switch for the tests that show the closed-form checks
 * catch a plausible error (0 = faithful oracle). */
static int g_mut = 0;
void orc_set_mutation(int m) { g_mut = m; }
int orc_mutation(void) { return g_mut; }
#define VPB 512  /* voxelPerBlock, GPUVoxelGrid.h:8 */

typedef struct { float x, y, z; } f3;
typedef struct { int x, y, z; } i3;

/* ------------------------------------------------------------------ */
/* scalar helpers                                                       */
/* ------------------------------------------------------------------ */

/* (int) of a float, as x86-64 cvttss2si does it: truncation toward zero,
 * INT_MIN for NaN and for values outside the int range.  The reference
 * writes plain `(int)` casts (GPUVoxelGrid.cpp:348-349, .h:136-187). */
int orc_f2i(float f)
{
    if (f >= -2147483648.0f && f < 2147483648.0f)
        return (int)f;
    return (int)0x80000000u;
}

/* getPrimeHash, GPUVoxelGrid.h:124-128: int arithmetic that wraps mod 2^32,
 * returned as unsigned long long (sign extension of the int result). */
uint64_t orc_prime_hash(int x, int y, int z)
{
    uint32_t h = (uint32_t)x * 73856093u + (uint32_t)y * 19349669u + (uint32_t)z * 83492791u;
    return (uint64_t)(int64_t)(int32_t)h;
}

/* getVBIndeces / getIndeces per axis, GPUVoxelGrid.h:131-190:
 * a >= 0 ? (int)(a / vbl) : (int)(a / vbl - 1). */
int orc_vb_index(float a, float vbl)
{
    if (g_mut == 3)
        return (int)floorf(a / vbl);
    if (a >= 0)
        return orc_f2i(a / vbl);
    return orc_f2i(a / vbl - 1.0f);
}

/* getVinVBIndeces, GPUVoxelGrid.h:192-221: p -= vbl * vbInd (float * int ->
 * float, kokkosFunctions.h:737-740), then the same split per axis with vl. */
void orc_vin_vb(const float p[3], const int vb[3], float vbl, float vl, int out[3])
{
    for (int a = 0; a < 3; a++) {
        float q = p[a] - vbl * (float)vb[a];
        out[a] = (q >= 0) ? orc_f2i(q / vl) : orc_f2i(q / vl - 1.0f);
    }
}

static inline i3 vb_of(f3 a, float vbl)
{
    i3 o = {orc_vb_index(a.x, vbl), orc_vb_index(a.y, vbl), orc_vb_index(a.z, vbl)};
    return o;
}

static inline i3 vin_of(f3 p, i3 vb, float vbl, float vl)
{
    float pp[3] = {p.x, p.y, p.z};
    int b[3] = {vb.x, vb.y, vb.z}, o[3];
    orc_vin_vb(pp, b, vbl, vl, o);
    i3 r = {o[0], o[1], o[2]};
    return r;
}

/* transformPoint, GPUVoxelGrid.h:240-247: row-major, summed left to right. */
static inline f3 xform(f3 p, const float *E)
{
    f3 o;
    o.x = p.x * E[0] + p.y * E[1] + p.z * E[2] + E[3];
    o.y = p.x * E[4] + p.y * E[5] + p.z * E[6] + E[7];
    o.z = p.x * E[8] + p.y * E[9] + p.z * E[10] + E[11];
    return o;
}

/* Kvec3f scalar product, kokkosFunctions.h:212-214. */
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* getVoxelMidPoint, GPUVoxelGrid.h:275-278: int*float in float, then
 * + (v + 0.5) * vl in double, rounded to float by the Kvec3f constructor. */
static inline f3 voxel_mid(i3 v, i3 vb, float vbl, float vl)
{
    f3 o;
    o.x = (float)((double)((float)vb.x * vbl) + ((double)v.x + 0.5) * (double)vl);
    o.y = (float)((double)((float)vb.y * vbl) + ((double)v.y + 0.5) * (double)vl);
    o.z = (float)((double)((float)vb.z * vbl) + ((double)v.z + 0.5) * (double)vl);
    return o;
}

/* Rigid pose preparation (stand-in for the Eigen float inverse used at
 * GPUVoxelGrid.cpp:55,109, which is not available here): E_inv = [R^T | -R^T t]
 * evaluated in double from the float entries and rounded once to float.
 * Ez = normalize(third column of E) in float, GPUVoxelGrid.cpp:52-53 with
 * Kvec3f::normalize, kokkosFunctions.h:169-180. */
void orc_pose_prepare(const float E[16], float Einv[16], float Ez[3])
{
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++)
            Einv[r * 4 + c] = (g_mut == 2) ? E[r * 4 + c] : E[c * 4 + r];
        double t = (double)E[0 * 4 + r] * (double)E[3] + (double)E[1 * 4 + r] * (double)E[7] +
                   (double)E[2 * 4 + r] * (double)E[11];
        Einv[r * 4 + 3] = (float)(-t);
    }
    Einv[12] = 0.0f; Einv[13] = 0.0f; Einv[14] = 0.0f; Einv[15] = 1.0f;
    float x = E[2], y = E[6], z = E[10];
    float len = sqrtf(x * x + y * y + z * z);
    Ez[0] = x / len;
    Ez[1] = y / len;
    Ez[2] = z / len;
}

/* Accepts a camera->world pose only if it is finite, has the homogeneous
 * last row the reference asserts (integrate.cpp:147-159) and an orthonormal
 * rotation block (|R^T R - I| <= 1e-4 entrywise). */
int orc_pose_is_rigid(const float E[16])
{
    for (int i = 0; i < 16; i++)
        if (!isfinite(E[i]))
            return 0;
    if (E[12] != 0.0f || E[13] != 0.0f || E[14] != 0.0f || E[15] != 1.0f)
        return 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0;
            for (int r = 0; r < 3; r++)
                s += (double)E[r * 4 + a] * (double)E[r * 4 + b];
            if (fabs(s - (a == b ? 1.0 : 0.0)) > 1e-4)
                return 0;
        }
    return 1;
}

/* ------------------------------------------------------------------ */
/* grid: hash (int32 key -> brick index) + brick arrays                 */
/* ------------------------------------------------------------------ */

struct orc_grid {
    float vl, vbl, trunc, rmin, dmax, wmax;
    int H, W, with_color;
    float fx, fy, cx, cy, ci, cj;
    int64_t max_bricks, n, cap;
    i3 *coord;
    int64_t *key;
    float *sdf, *w, *rgb;
    uint32_t *stamp;
    uint32_t epoch;
    /* open addressing over uint32 keys */
    uint64_t hcap;
    uint32_t *hkey;
    int64_t *hval; /* -1 = empty */
};

static inline uint32_t mix32(uint32_t h)
{
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

static int64_t h_find(const orc_grid *g, uint32_t k)
{
    uint64_t m = g->hcap - 1, b = mix32(k) & m;
    while (g->hval[b] >= 0) {
        if (g->hkey[b] == k)
            return g->hval[b];
        b = (b + 1) & m;
    }
    return -1;
}

static void h_put(orc_grid *g, uint32_t k, int64_t v)
{
    uint64_t m = g->hcap - 1, b = mix32(k) & m;
    while (g->hval[b] >= 0)
        b = (b + 1) & m;
    g->hkey[b] = k;
    g->hval[b] = v;
}

static void h_grow(orc_grid *g)
{
    uint64_t oc = g->hcap;
    uint32_t *ok = g->hkey;
    int64_t *ov = g->hval;
    g->hcap = oc * 2;
    g->hkey = (uint32_t *)malloc(g->hcap * sizeof(uint32_t));
    g->hval = (int64_t *)malloc(g->hcap * sizeof(int64_t));
    for (uint64_t i = 0; i < g->hcap; i++)
        g->hval[i] = -1;
    for (uint64_t i = 0; i < oc; i++)
        if (ov[i] >= 0)
            h_put(g, ok[i], ov[i]);
    free(ok);
    free(ov);
}

static void bricks_grow(orc_grid *g)
{
    int64_t nc = g->cap ? g->cap * 2 : 1024;
    g->coord = (i3 *)realloc(g->coord, nc * sizeof(i3));
    g->key = (int64_t *)realloc(g->key, nc * sizeof(int64_t));
    g->sdf = (float *)realloc(g->sdf, nc * VPB * sizeof(float));
    g->w = (float *)realloc(g->w, nc * VPB * sizeof(float));
    g->stamp = (uint32_t *)realloc(g->stamp, nc * sizeof(uint32_t));
    if (g->with_color)
        g->rgb = (float *)realloc(g->rgb, nc * VPB * 3 * sizeof(float));
    g->cap = nc;
}

orc_grid *orc_create(float voxel_length, float truncation, int height, int width,
                     int64_t max_bricks, int with_color, float max_weight,
                     float min_depth, float max_ray_length)
{
    orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
    g->vl = voxel_length;
    g->vbl = (float)VPE * voxel_length; /* GPUVoxelGrid.cpp:15 */
    g->trunc = truncation;
    g->H = height;
    g->W = width;
    g->max_bricks = max_bricks;
    g->with_color = with_color;
    g->wmax = max_weight;
    g->rmin = min_depth;
    g->dmax = max_ray_length;
    g->hcap = 1u << 16;
    g->hkey = (uint32_t *)malloc(g->hcap * sizeof(uint32_t));
    g->hval = (int64_t *)malloc(g->hcap * sizeof(int64_t));
    for (uint64_t i = 0; i < g->hcap; i++)
        g->hval[i] = -1;
    return g;
}

void orc_destroy(orc_grid *g)
{
    if (!g)
        return;
    free(g->coord); free(g->key); free(g->sdf); free(g->w); free(g->rgb);
    free(g->stamp); free(g->hkey); free(g->hval);
    free(g);
}

/* setIntrinsic, GPUVoxelGrid.cpp:253-266: ci = resi - cy, cj = cx. */
void orc_set_intrinsics(orc_grid *g, float fx, float fy, float cx, float cy)
{
    g->fx = fx;
    g->fy = fy;
    g->cx = cx;
    g->cy = cy;
    g->ci = (g_mut == 4) ? cy : (float)g->H - cy;
    g->cj = cx;
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* getPoint, GPUVoxelGrid.h:229-232. */
static inline f3 get_point(const orc_grid *g, int i, int j, float d)
{
    f3 p;
    p.x = ((float)j - g->cj) / g->fx * d;
    p.y = ((float)i - g->ci) / g->fy * d;
    p.z = d;
    return p;
}

/* depthToPcdTag, GPUVoxelGrid.h:303-307: (float)depth / 1000. */
static inline f3 input_pcd(const orc_grid *g, const uint16_t *depth, int pitch, int i, int j)
{
    return get_point(g, i, j, (float)depth[(int64_t)i * pitch + j] / 1000.0f);
}

/* A band's distinct blocks in order of first occurrence (collection). */
typedef struct {
    i3 *v;
    int64_t n, cap;
} band_list;

static void band_push(band_list *b, i3 c)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 4096;
        b->v = (i3 *)realloc(b->v, (size_t)b->cap * sizeof(i3));
    }
    b->v[b->n++] = c;
}

/* Open-addressing set of block coordinates (per collection thread). */
typedef struct {
    uint64_t *k; /* packed coordinate + 1; 0 = empty */
    uint64_t cap, n;
} bset;

static uint64_t bpack(i3 c)
{
    return (((uint64_t)(uint32_t)c.x & 0x1FFFFFu) << 42) | (((uint64_t)(uint32_t)c.y & 0x1FFFFFu) << 21) |
           ((uint64_t)(uint32_t)c.z & 0x1FFFFFu);
}

static void bset_init(bset *s)
{
    s->cap = 1u << 14;
    s->n = 0;
    s->k = (uint64_t *)calloc(s->cap, sizeof(uint64_t));
}

static void bset_free(bset *s) { free(s->k); }

static uint64_t bset_slot(const bset *s, uint64_t v)
{
    uint64_t h = v * 0x9E3779B97F4A7C15ull;
    return (h >> 20) & (s->cap - 1);
}

/* Adds c; returns 1 when it was not in the set.  Coordinates beyond 21 bits
 * per axis would collide when packed: they always return 1, so the band
 * lists every occurrence and the merge's table lookup drops the repeats. */
static int bset_add(bset *s, i3 c)
{
    if (c.x < -(1 << 20) || c.x >= (1 << 20) || c.y < -(1 << 20) || c.y >= (1 << 20) || c.z < -(1 << 20) ||
        c.z >= (1 << 20))
        return 1;
    const uint64_t v = bpack(c) + 1;
    uint64_t b = bset_slot(s, v);
    while (s->k[b]) {
        if (s->k[b] == v)
            return 0;
        b = (b + 1) & (s->cap - 1);
    }
    s->k[b] = v;
    if (++s->n * 2 > s->cap) {
        uint64_t oc = s->cap, *ok = s->k;
        s->cap *= 2;
        s->k = (uint64_t *)calloc(s->cap, sizeof(uint64_t));
        for (uint64_t i = 0; i < oc; i++)
            if (ok[i]) {
                uint64_t bb = bset_slot(s, ok[i]);
                while (s->k[bb])
                    bb = (bb + 1) & (s->cap - 1);
                s->k[bb] = ok[i];
            }
        free(ok);
    }
    return 1;
}

/* ------------------------------------------------------------------ */
/* integrate / deintegrate                                              */
/* ------------------------------------------------------------------ */

/* One voxel: updateSingleVoxel[Color] / deUpdateSingleVoxel[Color],
 * GPUVoxelGrid.cpp:345-469.  Returns 1 when the voxel was written. */
static int update_voxel(const orc_grid *g, float *s, float *w, float *rgb, f3 vm,
                        const uint16_t *depth, int pitch, const uint8_t *rgbin,
                        const float *E, const float *Einv, f3 Ez, int sign)
{
    f3 c = xform(vm, Einv);
    /* getCoordinate, GPUVoxelGrid.h:234-237 */
    float ci = c.y * g->fy / c.z + g->ci;
    float cj = c.x * g->fx / c.z + g->cj;
    int cori = orc_f2i(ci), corj = orc_f2i(cj);
    if (!(cori < g->H && corj < g->W && cori >= 0 && corj >= 0))
        return 0;
    f3 vpt = input_pcd(g, depth, pitch, cori, corj);
    if (!(vpt.z > g->rmin))
        return 0;
    vpt = xform(vpt, E);
    f3 dv = {vpt.x - vm.x, vpt.y - vm.y, vpt.z - vm.z};
    float sdf = dot3(dv, Ez);
    if (g_mut == 1)
        sdf = -sdf;
    if (!(fabsf(sdf) < g->trunc))
        return 0;
    float ww = *w;
    /* the colour of the pixel the voxel projects to (GPUVoxelGrid.cpp:406-416) */
    const uint8_t *px = rgb ? rgbin + (g_mut == 10 ? ((int64_t)corj * g->H + cori)
                                                   : ((int64_t)cori * g->W + corj)) * 3
                            : NULL;
    if (sign > 0) {
        *s = (*s * ww + sdf) / (ww + 1.0f);
        if (rgb)
            for (int ch = 0; ch < 3; ch++)
                rgb[ch] = (rgb[ch] * ww + (float)px[ch]) / (ww + (g_mut == 7 ? 2.0f : 1.0f));
        /* Project semantics (no counterpart in the reference, which never
         * caps, GPUVoxelGrid.h:14,62): with max_weight > 0 the weight stops
         * growing at the cap while the running average keeps w+1 as divisor. */
        if (g->wmax > 0.0f && (g_mut == 8 ? ww > g->wmax : ww >= g->wmax))
            *w = g->wmax;
        else
            *w = ww + 1.0f;
    } else {
        if (ww != 1.0f) {
            *s = (*s * ww - sdf) / (ww - 1.0f);
            if (rgb)
                for (int ch = 0; ch < 3; ch++)
                    rgb[ch] = (rgb[ch] * ww + (g_mut == 11 ? (float)px[ch] : -(float)px[ch])) / (ww - 1.0f);
            *w = ww - 1.0f;
        } else {
            *w = 0.0f;
            if (g_mut != 6)
                *s = 0.0f;
            if (rgb)
                rgb[0] = rgb[1] = rgb[2] = 0.0f;
        }
    }
    return 1;
}

int orc_integrate(orc_grid *g, const uint16_t *depth, int pitch, const uint8_t *rgb,
                  const float E[16], int sign, orc_counts *out)
{
    float Einv[16], Ezv[3];
    orc_pose_prepare(E, Einv, Ezv);
    f3 Ez = {Ezv[0], Ezv[1], Ezv[2]};
    g->epoch++;
    int overflow = 0;
    int64_t n_new = 0, n_valid = 0;
    int64_t tcap = 1024, tn = 0;
    int64_t *touched = (int64_t *)malloc(tcap * sizeof(int64_t));
    const float tr = g->trunc;

    /* vbCollectionTag, GPUVoxelGrid.cpp:471-514.  The allocated set is order
     * independent and brick indices are not observable.  Phase 1 (threads
     * over contiguous row bands) lists each band's distinct blocks in their
     * order of first occurrence; phase 2 inserts the bands' lists in band
     * order.  That is the serial pixel order's first-occurrence order, so
     * even the blocks a full pool refuses are those of a serial pass. */
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    if (nth > g->H)
        nth = g->H > 0 ? g->H : 1;
    band_list *bands = (band_list *)calloc((size_t)nth, sizeof(band_list));
#pragma omp parallel num_threads(nth) reduction(+ : n_valid)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        band_list *bl = &bands[t];
        const int r0 = (int)((int64_t)g->H * t / nth), r1 = (int)((int64_t)g->H * (t + 1) / nth);
        bset set;
        bset_init(&set);
        for (int i = r0; i < r1; i++) {
            for (int j = 0; j < g->W; j++) {
                f3 p = input_pcd(g, depth, pitch, i, j);
                if (!(p.z > g->rmin))
                    continue;
                n_valid++;
                f3 pt = xform(p, E);
                f3 lo = {pt.x - tr, pt.y - tr, pt.z - tr};
                f3 hi = {pt.x + tr, pt.y + tr, pt.z + tr};
                i3 mn = vb_of(lo, g->vbl), mx = vb_of(hi, g->vbl);
                for (int x = mn.x; x <= mx.x; x++)
                    for (int y = mn.y; y <= mx.y; y++)
                        for (int z = mn.z; z <= mx.z; z++) {
                            i3 c = {x, y, z};
                            if (bset_add(&set, c))
                                band_push(bl, c);
                        }
            }
        }
        bset_free(&set);
    }
    for (int t = 0; t < nth; t++) {
        for (int64_t e = 0; e < bands[t].n; e++) {
            const i3 c = bands[t].v[e];
            uint64_t key = orc_prime_hash(c.x, c.y, c.z);
            int64_t b = h_find(g, (uint32_t)key);
            if (b < 0) {
                if (g->n >= g->max_bricks) {
                    overflow = 1;
                    continue;
                }
                if (g->n == g->cap)
                    bricks_grow(g);
                b = g->n++;
                g->coord[b] = c;
                g->key[b] = (int64_t)key;
                memset(g->sdf + b * VPB, 0, VPB * sizeof(float));
                memset(g->w + b * VPB, 0, VPB * sizeof(float));
                if (g->with_color)
                    memset(g->rgb + b * VPB * 3, 0, VPB * 3 * sizeof(float));
                g->stamp[b] = 0;
                if ((uint64_t)g->n * 2 > g->hcap)
                    h_grow(g);
                h_put(g, (uint32_t)key, b);
                n_new++;
            }
            if (g->stamp[b] != g->epoch) {
                g->stamp[b] = g->epoch;
                if (tn == tcap) {
                    tcap *= 2;
                    touched = (int64_t *)realloc(touched, tcap * sizeof(int64_t));
                }
                touched[tn++] = b;
            }
        }
        free(bands[t].v);
    }
    free(bands);

    /* integrationCore[Color]Tag / deIntegrationCore[Color]Tag,
     * GPUVoxelGrid.cpp:516-550: vm = pzero + (k + 0.5) * vl in double. */
    int64_t n_upd = 0;
    const uint8_t *rgbin = (g->with_color && rgb) ? rgb : NULL;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : n_upd)
    for (int64_t t = 0; t < tn; t++) {
        int64_t b = touched[t];
        i3 c = g->coord[b];
        f3 pz = {g->vbl * (float)c.x, g->vbl * (float)c.y, g->vbl * (float)c.z}; /* .cpp:501 */
        for (int m = 0; m < VPE; m++)
            for (int l = 0; l < VPE; l++)
                for (int k = 0; k < VPE; k++) {
                    int v = k + VPE * l + VPE * VPE * m;
                    f3 vm;
                    vm.x = (float)((double)pz.x + ((double)k + 0.5) * (double)g->vl);
                    vm.y = (float)((double)pz.y + ((double)l + 0.5) * (double)g->vl);
                    vm.z = (float)((double)pz.z + ((double)m + 0.5) * (double)g->vl);
                    float *vr = rgbin ? g->rgb + (b * VPB + v) * 3 : NULL;
                    n_upd += update_voxel(g, g->sdf + b * VPB + v, g->w + b * VPB + v, vr, vm,
                                          depth, pitch, rgbin, E, Einv, Ez, sign);
                }
    }
    free(touched);
    if (out) {
        out->n_touched = tn;
        out->n_new = n_new;
        out->n_upd = n_upd;
        out->n_valid = n_valid;
    }
    return overflow ? 3 : 0;
}

/* ------------------------------------------------------------------ */
/* raycast                                                              */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t *vox_seen; /* n * 8 words */
    uint64_t kcap;
    uint32_t *kkey;
    uint8_t *kused;
    int64_t nkeys;
} ray_stats;

static void stats_key(ray_stats *st, uint32_t k)
{
    uint64_t m = st->kcap - 1, b = mix32(k) & m;
    while (st->kused[b]) {
        if (st->kkey[b] == k)
            return;
        b = (b + 1) & m;
    }
    st->kused[b] = 1;
    st->kkey[b] = k;
    st->nkeys++;
    if ((uint64_t)st->nkeys * 2 > st->kcap) {
        uint64_t oc = st->kcap;
        uint32_t *ok = st->kkey;
        uint8_t *ou = st->kused;
        st->kcap = oc * 2;
        st->kkey = (uint32_t *)calloc(st->kcap, sizeof(uint32_t));
        st->kused = (uint8_t *)calloc(st->kcap, 1);
        for (uint64_t i = 0; i < oc; i++)
            if (ou[i]) {
                uint64_t mm = st->kcap - 1, bb = mix32(ok[i]) & mm;
                while (st->kused[bb])
                    bb = (bb + 1) & mm;
                st->kused[bb] = 1;
                st->kkey[bb] = ok[i];
            }
        free(ok);
        free(ou);
    }
}

typedef struct {
    int64_t samples, probes;
} ray_local;

static inline int64_t probe(const orc_grid *g, i3 vb, ray_stats *st, ray_local *rl)
{
    uint32_t k = (uint32_t)orc_prime_hash(vb.x, vb.y, vb.z);
    rl->probes++;
    if (st)
        stats_key(st, k);
    return h_find(g, k);
}

static inline int vin_ok(i3 v)
{
    return v.x >= 0 && v.x < VPE && v.y >= 0 && v.y < VPE && v.z >= 0 && v.z < VPE;
}

/* raycastingTag, GPUVoxelGrid.cpp:561-715, for pixel (i, j).  Returns 1 on a
 * zero crossing, 0 on a miss, -flag on reference UB (output zeroed). */
static int march(const orc_grid *g, int i, int j, const float *Er, f3 *pout, f3 *cout,
                 ray_stats *st, ray_local *rl)
{
    const float vl = g->vl, vbl = g->vbl, dt = g->trunc, dmax = g->dmax;
    /* createNormPcdTag, GPUVoxelGrid.cpp:717-722 */
    f3 n = get_point(g, i, j, 1.0f);
    float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    n.x /= len; n.y /= len; n.z /= len;
    f3 rv = xform(n, Er);
    f3 dir = {rv.x - Er[3], rv.y - Er[7], rv.z - Er[11]};
    float l = 0.1f;
    float a = 1.0f - l;
    rv.x -= a * dir.x; rv.y -= a * dir.y; rv.z -= a * dir.z;

    int done = 0, Fn = 0, Lp = 0;
    i3 oldvb = {0, 0, 0};
    while (!done && !Fn && l < dmax) {
        i3 vb = vb_of(rv, vbl);
        int64_t idx = probe(g, vb, st, rl);
        int count = 0;
        while (idx < 0 && l < dmax) {
            rv.x += vbl * dir.x; rv.y += vbl * dir.y; rv.z += vbl * dir.z;
            l += vbl;
            vb = vb_of(rv, vbl);
            idx = probe(g, vb, st, rl);
            count++;
        }
        if (l >= dmax)
            break;
        if (count > 0) {
            float back = vbl - dt;
            rv.x -= back * dir.x; rv.y -= back * dir.y; rv.z -= back * dir.z;
            l -= back;
            vb = vb_of(rv, vbl);
            idx = probe(g, vb, st, rl);
            while (idx < 0 && l < dmax) {
                rv.x += dt * dir.x; rv.y += dt * dir.y; rv.z += dt * dir.z;
                l += dt;
                vb = vb_of(rv, vbl);
                idx = probe(g, vb, st, rl);
            }
        }
        if (l >= dmax)
            break;
        int64_t cur = idx;
        i3 vin = vin_of(rv, vb, vbl, vl);
        oldvb = vb;
        while (!done && l < dmax && oldvb.x == vb.x && oldvb.y == vb.y && oldvb.z == vb.z && !Fn) {
            if (!vin_ok(vin))
                return -ORC_UB_VOXEL_INDEX;
            int64_t cvi = cur * VPB + vin.x + VPE * vin.y + VPE * VPE * vin.z;
            rl->samples++;
            if (st)
                st->vox_seen[cvi >> 6] |= 1ull << (cvi & 63);
            float cw = g->w[cvi], cs = g->sdf[cvi];
            if (cw > 0) {
                if (cs > 0) {
                    Lp = 1;
                    rv.x += vl * dir.x; rv.y += vl * dir.y; rv.z += vl * dir.z;
                    l += vl;
                    oldvb = vb;
                    vb = vb_of(rv, vbl);
                    vin = vin_of(rv, vb, vbl, vl);
                } else if (Lp) {
                    done = 1;
                    f3 vm_in = voxel_mid(vin, vb, vbl, vl);
                    rv.x -= vl * dir.x; rv.y -= vl * dir.y; rv.z -= vl * dir.z;
                    l -= vl;
                    vb = vb_of(rv, vbl);
                    vin = vin_of(rv, vb, vbl, vl);
                    f3 vm_out = voxel_mid(vin, vb, vbl, vl);
                    int64_t ob = cur;
                    if (vb.x != oldvb.x || vb.y != oldvb.y || vb.z != oldvb.z) {
                        ob = probe(g, vb, st, rl);
                        if (ob < 0)
                            return -ORC_UB_MISSING_OUT;
                    }
                    if (!vin_ok(vin))
                        return -ORC_UB_VOXEL_INDEX;
                    int64_t ovi = ob * VPB + vin.x + VPE * vin.y + VPE * VPE * vin.z;
                    rl->samples++;
                    if (st)
                        st->vox_seen[ovi >> 6] |= 1ull << (ovi & 63);
                    float ow = g->w[ovi], os = g->sdf[ovi];
                    float x = (ow * os / vl + cw * (1.0f - cs / vl)) / (cw + ow);
                    if (g_mut == 5)
                        x = 1.0f - x;
                    f3 p;
                    p.x = vm_out.x + x * (vm_in.x - vm_out.x);
                    p.y = vm_out.y + x * (vm_in.y - vm_out.y);
                    p.z = vm_out.z + x * (vm_in.z - vm_out.z);
                    float s = dot3(p, dir) - dot3(rv, dir);
                    pout->x = rv.x + s * dir.x;
                    pout->y = rv.y + s * dir.y;
                    pout->z = rv.z + s * dir.z;
                    if (g->with_color) {
                        const float *co = g->rgb + ovi * 3, *cc = g->rgb + cvi * 3;
                        /* mutation 9 weights the inner voxel only */
                        const double wo = g_mut == 9 ? 0.0 : 0.5, wi = g_mut == 9 ? 1.0 : 0.5;
                        cout->x = (float)((wo * (double)co[0] + wi * (double)cc[0]) / 255.0);
                        cout->y = (float)((wo * (double)co[1] + wi * (double)cc[1]) / 255.0);
                        cout->z = (float)((wo * (double)co[2] + wi * (double)cc[2]) / 255.0);
                    }
                } else {
                    Fn = 1;
                }
            } else {
                Lp = 0;
                rv.x += vl * dir.x; rv.y += vl * dir.y; rv.z += vl * dir.z;
                l += vl;
                oldvb = vb;
                vb = vb_of(rv, vbl);
                vin = vin_of(rv, vb, vbl, vl);
            }
        }
    }
    return done;
}

void orc_raycast(orc_grid *g, const float E[16], float *points, float *colors, float *depth,
                 uint8_t *ub, int stats, orc_ray_counts *out)
{
    const int H = g->H, W = g->W;
    float Einv[16], Ezv[3];
    orc_pose_prepare(E, Einv, Ezv);
    int64_t n_hit = 0, n_ub = 0, samples = 0, probes = 0;
    ray_stats st;
    memset(&st, 0, sizeof(st));
    if (stats) {
        st.vox_seen = (uint64_t *)calloc((size_t)(g->n > 0 ? g->n : 1) * 8, sizeof(uint64_t));
        st.kcap = 1u << 16;
        st.kkey = (uint32_t *)calloc(st.kcap, sizeof(uint32_t));
        st.kused = (uint8_t *)calloc(st.kcap, 1);
    }
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : n_hit, n_ub, samples, probes) if (!stats)
    for (int i = 0; i < H; i++) {
        for (int j = 0; j < W; j++) {
            f3 p = {0, 0, 0}, c = {0, 0, 0};
            ray_local rl = {0, 0};
            int r = march(g, i, j, E, &p, &c, stats ? &st : NULL, &rl);
            samples += rl.samples;
            probes += rl.probes;
            uint8_t flag = 0;
            if (r < 0) {
                flag = (uint8_t)(-r);
                n_ub++;
                p.x = p.y = p.z = 0;
                c.x = c.y = c.z = 0;
            } else if (r > 0) {
                n_hit++;
            }
            int64_t o = (int64_t)i * W + j;
            points[o * 3 + 0] = p.x;
            points[o * 3 + 1] = p.y;
            points[o * 3 + 2] = p.z;
            if (colors) {
                colors[o * 3 + 0] = c.x;
                colors[o * 3 + 1] = c.y;
                colors[o * 3 + 2] = c.z;
            }
            if (depth) /* camera-frame z of the hit (project extension), 0 on a miss */
                depth[o] = (r > 0) ? (p.x * Einv[8] + p.y * Einv[9] + p.z * Einv[10] + Einv[11]) : 0.0f;
            if (ub)
                ub[o] = flag;
        }
    }
    if (out) {
        out->n_hit = n_hit;
        out->n_ub = n_ub;
        out->samples = samples;
        out->probes = probes;
        out->u_vox = 0;
        out->u_blk = 0;
        if (stats) {
            int64_t uv = 0;
            for (int64_t k = 0; k < g->n * 8; k++)
                uv += __builtin_popcountll(st.vox_seen[k]);
            out->u_vox = uv;
            out->u_blk = st.nkeys;
        }
    }
    if (stats) {
        free(st.vox_seen);
        free(st.kkey);
        free(st.kused);
    }
}

int64_t orc_n_bricks(const orc_grid *g) { return g->n; }

/* cleanMemory / cleanTag, GPUVoxelGrid.cpp:185-196, 785-805: every block
 * whose 512 weights are all zero is erased from the map.  Surviving bricks
 * keep their relative order (brick indices are not observable). */
int64_t orc_clean(orc_grid *g)
{
    int64_t kept = 0;
    for (int64_t b = 0; b < g->n; b++) {
        int empty = 1;
        for (int v = 0; v < VPB && empty; v++)
            if (g->w[b * VPB + v] != 0.0f)
                empty = 0;
        if (empty)
            continue;
        if (kept != b) {
            g->coord[kept] = g->coord[b];
            g->key[kept] = g->key[b];
            g->stamp[kept] = g->stamp[b];
            memmove(g->sdf + kept * VPB, g->sdf + b * VPB, VPB * sizeof(float));
            memmove(g->w + kept * VPB, g->w + b * VPB, VPB * sizeof(float));
            if (g->with_color)
                memmove(g->rgb + kept * VPB * 3, g->rgb + b * VPB * 3, VPB * 3 * sizeof(float));
        }
        kept++;
    }
    int64_t removed = g->n - kept;
    g->n = kept;
    for (uint64_t i = 0; i < g->hcap; i++)
        g->hval[i] = -1;
    for (int64_t b = 0; b < g->n; b++)
        h_put(g, (uint32_t)g->key[b], b);
    return removed;
}

/* Test hook: overwrites the colours of every brick (export order,
 * [n][512][3]), so a test can give voxels known colours before a raycast. */
void orc_import_rgb(orc_grid *g, const float *rgb)
{
    if (g->with_color && rgb)
        memcpy(g->rgb, rgb, (size_t)g->n * VPB * 3 * sizeof(float));
}

void orc_export(const orc_grid *g, int32_t *coords, int64_t *keys, float *sdf, float *w, float *rgb)
{
    for (int64_t b = 0; b < g->n; b++) {
        if (coords) {
            coords[b * 3 + 0] = g->coord[b].x;
            coords[b * 3 + 1] = g->coord[b].y;
            coords[b * 3 + 2] = g->coord[b].z;
        }
        if (keys)
            keys[b] = g->key[b];
    }
    if (sdf)
        memcpy(sdf, g->sdf, (size_t)g->n * VPB * sizeof(float));
    if (w)
        memcpy(w, g->w, (size_t)g->n * VPB * sizeof(float));
    if (rgb && g->with_color)
        memcpy(rgb, g->rgb, (size_t)g->n * VPB * 3 * sizeof(float));
}
