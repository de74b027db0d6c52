/*
 * icp_oracle.c -- CPU restatement of the reference's point maps, normal maps
 * and projective ICP sums.  TEST INFRASTRUCTURE ONLY (the checker; see
 * tsdf_oracle.h for who may load it).
 *
 * Follows, in binary32 arithmetic in source order (-ffp-contract=off):
 *   CreatePcdFromDepthView        src/utils/imageutil.cpp:89-111
 *   CalculateNormalsFromPCDView   src/utils/imageutil.cpp:116-157
 *   "Closest Point Kernel"        src/postprocessing/kokkosICP.cpp:69-115
 *   host sums of an iteration     src/postprocessing/kokkosICP.cpp:117-171
 * with the Kvec3f / Kmat4f arithmetic of src/VoxelGrid/kokkosFunctions.h
 * (:169-247 length, -, +, dot, cross; :690-700 Kmat4f * Kvec3f through a
 * Kvec4f (x, y, z, 1) dot product).  Maps are row-major [H][W][3].
 *
 * Planted mistakes (orc_set_mutation): 9 closes the search window at
 * centre + n; 10 takes the row of the projection from x instead of y; 11
 * swaps the operands of the first cross product of the normal.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "tsdf_oracle.h"

typedef struct {
    float x, y, z;
} v3;

static v3 v3sub(v3 a, v3 b)
{
    v3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}

static v3 v3add(v3 a, v3 b)
{
    v3 r = {a.x + b.x, a.y + b.y, a.z + b.z};
    return r;
}

/* Kvec3f::cross, kokkosFunctions.h:240-247 */
static v3 v3cross(v3 a, v3 b)
{
    v3 r;
    r.x = a.y * b.z - a.z * b.y;
    r.y = a.z * b.x - a.x * b.z;
    r.z = a.x * b.y - a.y * b.x;
    return r;
}

/* Kvec3f::length, kokkosFunctions.h:169-171 */
static float v3len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }

static v3 ld(const float *p, int64_t k)
{
    v3 r = {p[3 * k], p[3 * k + 1], p[3 * k + 2]};
    return r;
}

static void st(float *p, int64_t k, v3 v)
{
    p[3 * k] = v.x;
    p[3 * k + 1] = v.y;
    p[3 * k + 2] = v.z;
}

/* imageutil.cpp:89-111: ci = resi - cy, cj = cx (float); d = depth / 1000 */
void orc_depth_to_points(const uint16_t *depth, int H, int W, float fx, float fy, float cx, float cy, float *out)
{
    const float ci = (float)H - cy, cj = cx;
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            const float d = (float)depth[(int64_t)i * W + j] / 1000.0f;
            v3 p = {((float)j - cj) / fx * d, ((float)i - ci) / fy * d, d};
            st(out, (int64_t)i * W + j, p);
        }
}

/* imageutil.cpp:116-157 */
void orc_normals(const float *pcd, int H, int W, float *out)
{
    for (int64_t k = 0; k < (int64_t)H * W; k++) {
        v3 z = {0.0f, 0.0f, 1.0f};
        st(out, k, z);
    }
    for (int i = 1; i < H - 1; i++)
        for (int j = 1; j < W - 1; j++) {
            const int64_t k = (int64_t)i * W + j;
            const v3 p = ld(pcd, k), pd = ld(pcd, k + W), pu = ld(pcd, k - W), pr = ld(pcd, k + 1),
                     pl = ld(pcd, k - 1);
            if (p.z != 0 && pd.z != 0 && pu.z != 0 && pr.z != 0 && pl.z != 0) {
                const v3 n1 = orc_mutation() == 11 ? v3cross(v3sub(pr, p), v3sub(pd, p))
                                                   : v3cross(v3sub(pd, p), v3sub(pr, p));
                const v3 n2 = v3cross(v3sub(pl, p), v3sub(pd, p));
                const v3 n3 = v3cross(v3sub(pr, p), v3sub(pu, p));
                const v3 n4 = v3cross(v3sub(pu, p), v3sub(pl, p));
                v3 s = v3add(v3add(v3add(n1, n2), n3), n4);
                const float l = v3len(s);
                s.x /= l;
                s.y /= l;
                s.z /= l;
                st(out, k, s);
            }
        }
}

/* kokkosICP.cpp:69-115.  T: row-major 4x4 float (the float cast of Txy). */
void orc_icp_associate(const float *x, const float *y, const float *yn, const float *T, int H, int W, float fx,
                       float fy, float cx, float cy, int type, int n_search, float max_dist, float *out)
{
    /* kokkosICP.cpp:44-47: ci = resi - cy (rounded to float), cj = cx */
    const float ci_ = (float)((double)H - (double)cy), cj_ = cx;
    const int up = orc_mutation() == 9 ? 1 : 0;
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            const int64_t k = (int64_t)i * W + j;
            v3 best = {0.0f, 0.0f, 0.0f};
            const v3 xp = ld(x, k);
            if (xp.z != 0) {
                /* Kmat4f * Kvec3f: each row dotted with (x, y, z, 1) */
                v3 px;
                px.x = T[0] * xp.x + T[1] * xp.y + T[2] * xp.z + T[3] * 1.0f;
                px.y = T[4] * xp.x + T[5] * xp.y + T[6] * xp.z + T[7] * 1.0f;
                px.z = T[8] * xp.x + T[9] * xp.y + T[10] * xp.z + T[11] * 1.0f;
                const float c_x = (orc_mutation() == 10 ? px.x * fy : px.y * fy) / px.z + ci_;
                const float c_y = px.x * fx / px.z + cj_;
                const int centeri = orc_f2i(c_x), centerj = orc_f2i(c_y);
                float smallest = 1e6f;
                if (centeri > n_search && centeri < H - n_search && centerj > n_search && centerj < W - n_search) {
                    for (int r = centeri - n_search; r < centeri + n_search + up; r++)
                        for (int c = centerj - n_search; c < centerj + n_search + up; c++) {
                            const int64_t q = (int64_t)r * W + c;
                            const v3 py = ld(y, q);
                            if (type == 0) {
                                if (py.z != 0) {
                                    const float dist = v3len(v3sub(py, px));
                                    if (dist < smallest && dist < max_dist) {
                                        smallest = dist;
                                        best = py;
                                    }
                                }
                            } else {
                                const v3 nv = ld(yn, q);
                                if (py.z != 0 && nv.x != 0) {
                                    const v3 dv = v3sub(py, px);
                                    const float d = dv.x * nv.x + dv.y * nv.y + dv.z * nv.z;
                                    const float dist = fabsf(d);
                                    if (dist < smallest && dist < max_dist) {
                                        smallest = dist;
                                        v3 dn = {d * nv.x, d * nv.y, d * nv.z};
                                        best = v3add(px, dn);
                                    }
                                }
                            }
                        }
                }
            }
            st(out, k, best);
        }
}

/* kokkosICP.cpp:117-171 in the reference's loop order: the first pass sums
 * the centroids (double), the RMSE (float, of the float-rounded double norm)
 * and the counts (float); the second pass the centred cross-covariance H.
 * T: row-major 4x4 double (Txy_eigen).  out: [0] points of x, [1]
 * correspondences, [2..4] centre of x, [5..7] centre of x_nearest, [8..16]
 * H row-major, [17] RMSE (sqrtf of the float mean). */
void orc_icp_sums(const float *x, const float *xn, const double *T, int H, int W, double *out)
{
    double cxs[3] = {0, 0, 0}, cys[3] = {0, 0, 0};
    float rsme = 0.0f, n_cors = 0.0f, n_pts = 0.0f;
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            const int64_t k = (int64_t)i * W + j;
            const v3 a = ld(x, k), b = ld(xn, k);
            if (b.z != 0) {
                cxs[0] += a.x;
                cxs[1] += a.y;
                cxs[2] += a.z;
                cys[0] += b.x;
                cys[1] += b.y;
                cys[2] += b.z;
                n_cors++;
                const double av[4] = {a.x, a.y, a.z, 1.0};
                double e2 = 0.0;
                for (int r = 0; r < 4; r++) {
                    double t = 0.0;
                    for (int c = 0; c < 4; c++)
                        t += T[4 * r + c] * av[c];
                    const double bv = r == 0 ? b.x : r == 1 ? b.y : r == 2 ? b.z : 1.0;
                    e2 += (bv - t) * (bv - t);
                }
                const float tmp = (float)sqrt(e2);
                rsme += tmp * tmp;
            }
            if (a.z != 0)
                n_pts++;
        }
    double cx[3], cy[3];
    for (int r = 0; r < 3; r++) {
        cx[r] = cxs[r] / n_cors;
        cy[r] = cys[r] / n_cors;
    }
    double Hm[9] = {0};
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            const int64_t k = (int64_t)i * W + j;
            const v3 a = ld(x, k), b = ld(xn, k);
            if (a.z != 0 && b.z != 0) {
                const double xa[3] = {a.x - cx[0], a.y - cx[1], a.z - cx[2]};
                const double yb[3] = {b.x - cy[0], b.y - cy[1], b.z - cy[2]};
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++)
                        Hm[3 * r + c] += xa[r] * yb[c];
            }
        }
    rsme /= n_cors;
    rsme = sqrtf(rsme);
    out[0] = n_pts;
    out[1] = n_cors;
    for (int r = 0; r < 3; r++) {
        out[2 + r] = cx[r];
        out[5 + r] = cy[r];
    }
    memcpy(out + 8, Hm, sizeof(Hm));
    out[17] = rsme;
}
