// synth.cpp -- seeded synthetic depth frames for tests and benchmarks.
//
// Renders uint16 depth (mm) of an analytic scene by casting one ray per
// pixel in double precision with the reference camera convention
// (GPUVoxelGrid.h:229-232 with ci = H - cy, cj = cx, GPUVoxelGrid.cpp:259-260):
// the camera ray of pixel (i, j) is ((j - cj) / fx, (i - ci) / fy, 1), so the
// hit parameter is the camera-frame z.  Gaussian noise and tile dropout come
// from a counter-based generator (splitmix64 of seed and pixel index), so the
// frames are identical on every machine and thread count.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

struct Prim {
    int32_t kind;  // 0 plane (n, d) | 1 sphere (c, r) | 2 solid box (min, max) | 3 room interior (min, max)
    int32_t pad;
    double a[6];
};

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline double u01(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// smallest t > eps of the ray o + t d against one primitive, or +inf
double hit(const Prim& p, const double* o, const double* d)
{
    const double eps = 1e-9, inf = INFINITY;
    switch (p.kind) {
    case 0: {
        const double den = p.a[0] * d[0] + p.a[1] * d[1] + p.a[2] * d[2];
        if (std::fabs(den) < 1e-15)
            return inf;
        const double t = (p.a[3] - (p.a[0] * o[0] + p.a[1] * o[1] + p.a[2] * o[2])) / den;
        return t > eps ? t : inf;
    }
    case 1: {
        const double oc[3] = {o[0] - p.a[0], o[1] - p.a[1], o[2] - p.a[2]};
        const double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
        const double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - p.a[3] * p.a[3];
        const double disc = B * B - A * C;
        if (disc < 0)
            return inf;
        const double sq = std::sqrt(disc);
        double t = (-B - sq) / A;
        if (t > eps)
            return t;
        t = (-B + sq) / A;
        return t > eps ? t : inf;
    }
    case 2:
    case 3: {
        double tmin = -inf, tmax = inf;
        for (int k = 0; k < 3; ++k) {
            if (std::fabs(d[k]) < 1e-15) {
                if (o[k] < p.a[k] || o[k] > p.a[k + 3])
                    return inf;
                continue;
            }
            double t0 = (p.a[k] - o[k]) / d[k], t1 = (p.a[k + 3] - o[k]) / d[k];
            if (t0 > t1) {
                const double tt = t0;
                t0 = t1;
                t1 = tt;
            }
            tmin = t0 > tmin ? t0 : tmin;
            tmax = t1 < tmax ? t1 : tmax;
        }
        if (tmin > tmax)
            return inf;
        if (p.kind == 2)
            return tmin > eps ? tmin : inf;  // solid: entry point; origin inside sees nothing
        return tmax > eps ? tmax : inf;  // room: where the ray leaves the box
    }
    }
    return inf;
}

}  // namespace

extern "C" {

// pose: row-major camera->world 4x4 (double).  keep: fraction of drop tiles
// kept (1 = no dropout), tile: dropout tile edge in pixels.
int syn_render_depth(const void* prims_v, int n_prims, const double* pose, int H, int W, double fx, double fy,
                     double cx, double cy, double noise_mm, double max_range_m, uint64_t seed, double keep, int tile,
                     uint16_t* out)
{
    const Prim* prims = (const Prim*)prims_v;
    if (!prims || n_prims < 0 || !pose || H <= 0 || W <= 0 || !out || tile <= 0)
        return 1;
    const double ci = (double)H - cy, cj = cx;
    const double o[3] = {pose[3], pose[7], pose[11]};
    const uint64_t drop_seed = splitmix64(seed ^ 0xD1B54A32D192ED03ull);
    const int tiles_x = (W + tile - 1) / tile;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < H; ++i) {
        for (int j = 0; j < W; ++j) {
            const double dc[3] = {((double)j - cj) / fx, ((double)i - ci) / fy, 1.0};
            const double d[3] = {pose[0] * dc[0] + pose[1] * dc[1] + pose[2] * dc[2],
                                 pose[4] * dc[0] + pose[5] * dc[1] + pose[6] * dc[2],
                                 pose[8] * dc[0] + pose[9] * dc[1] + pose[10] * dc[2]};
            double t = INFINITY;
            for (int k = 0; k < n_prims; ++k) {
                const double h = hit(prims[k], o, d);
                t = h < t ? h : t;
            }
            uint16_t v = 0;
            if (std::isfinite(t) && t <= max_range_m) {
                double mm = t * 1000.0;
                if (noise_mm > 0) {
                    const uint64_t pix = (uint64_t)i * (uint64_t)W + (uint64_t)j;
                    const uint64_t h1 = splitmix64(seed ^ (pix * 0x9E3779B97F4A7C15ull));
                    const uint64_t h2 = splitmix64(h1);
                    const double g = std::sqrt(-2.0 * std::log(u01(h1))) * std::cos(6.283185307179586 * u01(h2));
                    mm += noise_mm * g;
                }
                const double r = std::nearbyint(mm);
                if (r >= 1.0 && r <= 65535.0)
                    v = (uint16_t)r;
            }
            if (keep < 1.0) {
                const uint64_t tid = (uint64_t)(i / tile) * (uint64_t)tiles_x + (uint64_t)(j / tile);
                if (u01(splitmix64(drop_seed ^ (tid * 0xC2B2AE3D27D4EB4Full))) >= keep)
                    v = 0;
            }
            out[(size_t)i * W + j] = v;
        }
    }
    return 0;
}

}  // extern "C"
