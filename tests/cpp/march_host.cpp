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
    const int mode = use_sb;  // empty-space skip: 0 none, 1 superblock distance field, 2 tile depth bounds
    use_sb = mode == 1;
    rp.use_sb = 0;
    rp.use_tb = 0;
    if (mode) {  // the certificate the raycast kernel evaluates
        const int reach = (int)(rp.dmax / rp.vbl) + 3;
        const int c[3] = {vb_index(E[3], rp.vbl), vb_index(E[7], rp.vbl), vb_index(E[11], rp.vbl)};
        int plo[3], phi[3];
        for (int a = 0; a < 3; ++a) {
            plo[a] = c[a] - reach;
            phi[a] = c[a] + reach;
        }
        rp.use_sb = (use_sb && alias_free(bmin, bmax, plo, phi)) ? 1 : 0;
        rp.use_tb = (mode == 2 && alias_free(bmin, bmax, plo, phi)) ? 1 : 0;
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
    // tile depth bounds, as k_ray_bounds builds them (grow as the host side
    // of tsdf_raycast computes it for this pose)
    const int tx = (W + kTbTile - 1) / kTbTile, ty = (H + kTbTile - 1) / kTbTile;
    std::vector<unsigned long long> tbz((size_t)tx * ty, ~0ull);
    int near = 0;
    if (rp.use_tb) {
        const double vl_ = vl, dt_ = trunc, vbl_ = g.vbl, dm = dmax;
        const double n_upd = 4.0 * (dm + vbl_) / (vl_ < dt_ ? vl_ : dt_) + 64.0;
        double tmax = 0.0;
        for (int a = 0; a < 3; ++a)
            tmax = std::fmax(tmax, std::fabs((double)E[4 * a + 3]));
        auto spacing = [](double x) { return std::ldexp(1.0, std::ilogb(x) - 23); };
        const double R = tmax + dm + vbl_ + 1.0;
        const float grow = (float)(std::sqrt(3.0) * n_upd * spacing(R) / 2.0 + n_upd * spacing(dm + vbl_) / 2.0 +
                                   std::sqrt(3.0) * n_upd * spacing(vbl_) / 2.0 + std::sqrt(3.0) * dm * spacing(R) +
                                   8.0 * spacing(R) + 1e-4);
        const double ex = std::fmax(std::fabs((0.0 - g.cam.cj) / g.cam.fx), std::fabs((W - 1.0 - g.cam.cj) / g.cam.fx));
        const double ey = std::fmax(std::fabs((0.0 - g.cam.ci) / g.cam.fy), std::fabs((H - 1.0 - g.cam.ci) / g.cam.fy));
        const float z_lo = (float)(0.1 / std::sqrt(1.0 + ex * ex + ey * ey) * (1.0 - 1e-4) - grow);
        near = !(z_lo > 1e-3f);
        for (int b = 0; b < n && !near; ++b) {
            const int* c = coords + 3 * b;
            float zmin;
            int r0, r1, c0, c1;
            if (!brick_bounds(c[0], c[1], c[2], g.vbl, P, g.cam, H, W, grow, z_lo, zmin, r0, r1, c0, c1))
                continue;
            const unsigned long long v = tb_pack(1u, zmin);
            for (int y = r0 / kTbTile; y <= r1 / kTbTile; ++y)
                for (int x = c0 / kTbTile; x <= c1 / kTbTile; ++x)
                    if (tbz[(size_t)y * tx + x] > v)
                        tbz[(size_t)y * tx + x] = v;
        }
        if (near)
            rp.use_tb = 0;
    }
    *certified = mode == 1 ? rp.use_sb : (mode == 2 ? rp.use_tb : 0);
    Footprint fp;
    std::memset(&fp, 0, sizeof(fp));
    uint64_t ns = 0, np = 0;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            const float zt = rp.use_tb ? tb_depth(tbz[(size_t)(i / kTbTile) * tx + j / kTbTile], 1u) : -1.0f;
            const RayResult r = march<float, false, false>(g, P, rp, i, j, fp, zt);
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
