// Host build of the raycast march (superscanner_amd/csrc/tsdf_march.hpp) for
// tests/test_march_host.py: the same source the gfx950 kernel runs, executed
// on the CPU over a grid rebuilt from the oracle's export, so its results can
// be compared with the oracle's raycast bit for bit without a GPU.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <climits>
#include <vector>

#include "../../superscanner_amd/csrc/tsdf_march.hpp"

using namespace tsdf;

extern "C" int march_host(int H, int W, float fx, float fy, float cx, float cy, float vl, float trunc, float dmax,
                          int n, const int64_t* keys, const int32_t* coords, const float* sdf, const float* w,
                          uint32_t n_buckets, int use_sb, const float* E, float* points, float* depth, uint8_t* ub,
                          uint64_t* samples, uint64_t* probes, int* certified)
{
    if (n_buckets == 0 || (n_buckets & (n_buckets - 1)) || (uint64_t)n * 2 > n_buckets)
        return 1;
    std::vector<unsigned long long> slots(n_buckets, kSlotEmpty);
    const uint32_t mask = n_buckets - 1;
    for (int b = 0; b < n; ++b) {
        const uint32_t key = (uint32_t)keys[b];
        uint32_t s = mix32(key) & mask;
        while (slots[s] != kSlotEmpty)
            s = (s + 1) & mask;
        slots[s] = ((unsigned long long)key << 32) | (uint32_t)b;
    }
    // superblock bits and bounding box of the inserted blocks, as the
    // allocation kernel maintains them
    std::vector<uint32_t> sb(kSbWords, 0u);
    int bmin[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, bmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (int b = 0; b < n; ++b) {
        const int* c = coords + 3 * b;
        const int k = sb_index(c[0], c[1], c[2]);
        if (k >= 0)
            sb[k >> 5] |= 1u << (k & 31);
        for (int a = 0; a < 3; ++a) {
            bmin[a] = c[a] < bmin[a] ? c[a] : bmin[a];
            bmax[a] = c[a] > bmax[a] ? c[a] : bmax[a];
        }
    }
    std::vector<float> pay((size_t)n * 1024);
    for (size_t v = 0; v < (size_t)n * 512; ++v) {
        pay[2 * v] = sdf[v];
        pay[2 * v + 1] = w[v];
    }
    GridView g;
    std::memset(&g, 0, sizeof(g));
    g.slots = slots.data();
    g.sb_bits = sb.data();
    g.payload = pay.data();
    g.mask = mask;
    g.n_bricks = (uint32_t)n;
    g.vl = vl;
    g.vbl = 8.0f * vl;
    g.trunc = trunc;
    g.dmax = dmax;
    g.H = H;
    g.W = W;
    g.cam.fx = fx;
    g.cam.fy = fy;
    g.cam.ci = (float)H - cy;
    g.cam.cj = cx;
    Pose P;
    std::memset(&P, 0, sizeof(P));
    for (int k = 0; k < 12; ++k)
        P.E[k] = E[k];
    for (int r = 0; r < 3; ++r) {  // rigid inverse in double, rounded once (as tsdf_pose_prepare)
        for (int c = 0; c < 3; ++c)
            P.Ei[r * 4 + c] = E[c * 4 + r];
        const double t = (double)E[r] * (double)E[3] + (double)E[4 + r] * (double)E[7] +
                         (double)E[8 + r] * (double)E[11];
        P.Ei[r * 4 + 3] = (float)(-t);
    }
    RayParams rp;
    rp.vl = vl;
    rp.vbl = g.vbl;
    rp.dt = trunc;
    rp.dmax = dmax;
    rp.Rvl = 1.0 / (double)vl;
    rp.Rvbl = 1.0 / (double)rp.vbl;
    rp.back = rp.vbl - rp.dt;
    rp.H = H;
    rp.W = W;
    rp.xcd_map = 0;
    rp.use_sb = 0;
    if (use_sb) {  // the certificate the raycast kernel evaluates
        const int reach = (int)(rp.dmax / rp.vbl) + 3;
        const int c[3] = {vb_index(E[3], rp.vbl), vb_index(E[7], rp.vbl), vb_index(E[11], rp.vbl)};
        int plo[3], phi[3];
        for (int a = 0; a < 3; ++a) {
            plo[a] = c[a] - reach;
            phi[a] = c[a] + reach;
        }
        rp.use_sb = alias_free(bmin, bmax, plo, phi) ? 1 : 0;
    }
    std::vector<uint8_t> df, df_tmp;
    rp.df = nullptr;
    for (int a = 0; a < 3; ++a)
        rp.df_lo[a] = rp.df_dim[a] = 0;
    if (rp.use_sb && n > 0) {  // the distance field of the superblock bits (as k_sbdf builds it)
        size_t vol = 1;
        for (int a = 0; a < 3; ++a) {
            rp.df_lo[a] = bmin[a] >> 3;
            rp.df_dim[a] = (bmax[a] >> 3) - rp.df_lo[a] + 1;
            vol *= (size_t)rp.df_dim[a];
        }
        df.resize(vol);
        df_tmp.resize(vol);
        sbdf_build_host(sb.data(), rp.df_lo, rp.df_dim, df.data(), df_tmp.data());
        rp.df = df.data();
    }
    *certified = rp.use_sb;
    Footprint fp;
    std::memset(&fp, 0, sizeof(fp));
    uint64_t ns = 0, np = 0;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            const RayResult r = march<float, false, false>(g, P, rp, i, j, fp);
            const size_t o = (size_t)i * W + j;
            points[o * 3 + 0] = r.p.x;
            points[o * 3 + 1] = r.p.y;
            points[o * 3 + 2] = r.p.z;
            depth[o] = r.code == 1 ? (r.p.x * P.Ei[8] + r.p.y * P.Ei[9] + r.p.z * P.Ei[10] + P.Ei[11]) : 0.0f;
            ub[o] = (uint8_t)(r.code < 0 ? -r.code : 0);
            ns += r.samples;
            np += r.probes;
        }
    *samples = ns;
    *probes = np;
    return 0;
}
