// Host build of the raycast march (superscanner_amd/csrc/tsdf_march.hpp) for
// tests/test_march_host.py: the same source the gfx950 kernel runs, executed
// on the CPU over a grid rebuilt from the oracle's export, so its results can
// be compared with the oracle's raycast bit for bit without a GPU.
// march_tile_reuse replays one raycast tile by tile as the kernel groups its
// rays (16 x 16 pixels a workgroup, four 8 x 8 wavefronts) and counts how
// often a tile's loaded hash lookups and voxel reads repeat: what a
// per-workgroup LDS cache of either could serve.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <climits>
#include <unordered_set>
#include <vector>

#include "../../superscanner_amd/csrc/tsdf_march.hpp"

using namespace tsdf;

namespace {

// The grid, pose and ray parameters of one raycast, as the kernel sees them.
struct HostRaycast {
    std::vector<unsigned long long> slots;
    std::vector<uint32_t> sb;
    std::vector<float> pay;
    std::vector<uint8_t> df, df_tmp;
    std::vector<unsigned long long> tbz;
    GridView g;
    Pose P;
    RayParams rp;
    int tx = 0;
    int certified = 0;

    float tile_depth(int i, int j) const
    {
        return rp.use_tb ? tb_depth(tbz[(size_t)(i / kTbTile) * tx + j / kTbTile], 1u) : -1.0f;
    }
};

int build(HostRaycast& h, int H, int W, float fx, float fy, float cx, float cy, float vl, float trunc, float dmax,
          int n, const int64_t* keys, const int32_t* coords, const float* sdf, const float* w, uint32_t n_buckets,
          int use_sb, const float* E)
{
    if (n_buckets == 0 || (n_buckets & (n_buckets - 1)) || (uint64_t)n * 2 > n_buckets)
        return 1;
    std::vector<unsigned long long>& slots = h.slots;
    slots.assign(n_buckets, kSlotEmpty);
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
    std::vector<uint32_t>& sb = h.sb;
    sb.assign(kSbWords, 0u);
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
    std::vector<float>& pay = h.pay;
    pay.resize((size_t)n * 1024);
    for (size_t v = 0; v < (size_t)n * 512; ++v) {
        pay[2 * v] = sdf[v];
        pay[2 * v + 1] = w[v];
    }
    GridView& g = h.g;
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
    Pose& P = h.P;
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
    RayParams& rp = h.rp;
    std::memset(&rp, 0, sizeof(rp));
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
    std::vector<uint8_t>&df = h.df, &df_tmp = h.df_tmp;
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
    h.tx = tx;
    std::vector<unsigned long long>& tbz = h.tbz;
    tbz.assign((size_t)tx * ty, ~0ull);
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
    h.certified = mode == 1 ? rp.use_sb : (mode == 2 ? rp.use_tb : 0);
    return 0;
}

}  // namespace

extern "C" int march_host(int H, int W, float fx, float fy, float cx, float cy, float vl, float trunc, float dmax,
                          int n, const int64_t* keys, const int32_t* coords, const float* sdf, const float* w,
                          uint32_t n_buckets, int use_sb, const float* E, float* points, float* depth, uint8_t* ub,
                          uint64_t* samples, uint64_t* probes, int* certified)
{
    HostRaycast h;
    if (build(h, H, W, fx, fy, cx, cy, vl, trunc, dmax, n, keys, coords, sdf, w, n_buckets, use_sb, E))
        return 1;
    const GridView& g = h.g;
    const Pose& P = h.P;
    const RayParams& rp = h.rp;
    *certified = h.certified;
    Footprint fp;
    std::memset(&fp, 0, sizeof(fp));
    uint64_t ns = 0, np = 0;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            const RayResult r = march<float, false, false>(g, P, rp, i, j, fp, h.tile_depth(i, j));
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

namespace {

struct TileTrace {
    std::vector<uint32_t> ray_keys[256];
    int ray = 0;
    uint64_t voxel_reads = 0;
    std::unordered_set<uint64_t> voxels;
};

void on_key(void* c, uint32_t key)
{
    TileTrace* t = (TileTrace*)c;
    t->ray_keys[t->ray].push_back(key);
}

void on_voxel(void* c, int id, int v)
{
    TileTrace* t = (TileTrace*)c;
    ++t->voxel_reads;
    t->voxels.insert(((uint64_t)(uint32_t)id << 9) | (uint64_t)v);
}

}  // namespace

// stats: [0] loaded lookups, [1] distinct (tile, key) pairs, [2] voxel reads,
// [3] distinct (tile, voxel) pairs, [4] lane lookups hitting in lockstep,
// [5] wavefront lookup steps, [6] of them with every lane hitting, [7] the
// most distinct keys of one tile, [8] samples, [9] probes.
// Lockstep: lookup number p of every ray of a tile is taken at the same time
// by its wavefront; it hits if its key was looked up at an earlier number by
// any ray of the tile (the best a cache filled by the tile's own misses can
// do while a wavefront waits for its slowest lane).
extern "C" int march_tile_reuse(int H, int W, float fx, float fy, float cx, float cy, float vl, float trunc,
                                float dmax, int n, const int64_t* keys, const int32_t* coords, const float* sdf,
                                const float* w, uint32_t n_buckets, int use_sb, const float* E, uint64_t* stats)
{
    HostRaycast h;
    if (build(h, H, W, fx, fy, cx, cy, vl, trunc, dmax, n, keys, coords, sdf, w, n_buckets, use_sb, E))
        return 1;
    Footprint fp;
    std::memset(&fp, 0, sizeof(fp));
    TileTrace t;
    HostTrace hook = {on_key, on_voxel, &t};
    host_trace() = &hook;
    for (int k = 0; k < 10; ++k)
        stats[k] = 0;
    for (int ti = 0; ti < H; ti += 16)
        for (int tj = 0; tj < W; tj += 16) {
            for (auto& v : t.ray_keys)
                v.clear();
            t.voxels.clear();
            t.voxel_reads = 0;
            for (int i = ti; i < ti + 16 && i < H; ++i)
                for (int j = tj; j < tj + 16 && j < W; ++j) {
                    t.ray = (i - ti) * 16 + (j - tj);
                    const RayResult r = march<float, false, true>(h.g, h.P, h.rp, i, j, fp, h.tile_depth(i, j));
                    stats[8] += r.samples;
                    stats[9] += r.probes;
                }
            std::unordered_set<uint32_t> tile_keys, seen;
            size_t longest = 0;
            for (const auto& v : t.ray_keys) {
                stats[0] += v.size();
                tile_keys.insert(v.begin(), v.end());
                longest = v.size() > longest ? v.size() : longest;
            }
            stats[1] += tile_keys.size();
            stats[7] = tile_keys.size() > stats[7] ? tile_keys.size() : stats[7];
            stats[2] += t.voxel_reads;
            stats[3] += t.voxels.size();
            for (size_t p = 0; p < longest; ++p) {
                for (int wv = 0; wv < 4; ++wv) {
                    bool any = false, all = true;
                    for (int ln = 0; ln < 64; ++ln) {
                        const auto& v = t.ray_keys[((wv >> 1) * 8 + (ln >> 3)) * 16 + (wv & 1) * 8 + (ln & 7)];
                        if (p < v.size()) {
                            any = true;
                            if (seen.count(v[p]))
                                ++stats[4];
                            else
                                all = false;
                        }
                    }
                    stats[5] += any ? 1 : 0;
                    stats[6] += (any && all) ? 1 : 0;
                }
                for (const auto& v : t.ray_keys)
                    if (p < v.size())
                        seen.insert(v[p]);
            }
        }
    host_trace() = nullptr;
    return 0;
}
