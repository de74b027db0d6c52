// tsdf_march.hpp -- the raycast march of the TSDF engine and the device view
// of the grid it reads.
//
// The march is __host__ __device__: the HIP kernel runs it on gfx950, and
// tests/cpp/test_march_host.cpp runs the very same code on the CPU against
// the oracle.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsdf_math.hpp"

#pragma clang fp contract(off)

namespace tsdf {

// Device-side counters of one grid.
struct Ctrl {
    unsigned int alloc_top;         // brick ids handed out (may pass n_bricks)
    unsigned int overflow;          // refused allocations / full table
    unsigned int touched_count[2];  // touched-list length, by call parity
    unsigned long long total_touched;
    unsigned long long total_new;
    unsigned long long total_updated;
    unsigned long long total_valid;
    unsigned long long ray_hits;
    unsigned long long ray_ub;
    unsigned long long ray_samples;
    unsigned long long ray_probes;
    int bbox_min[3];  // bounding box of every block ever inserted since reset
    int bbox_max[3];
};

// Device view of one grid (all pointers are device memory on the GPU).
struct GridView {
    unsigned long long* slots;  // [n_buckets] (key << 32) | brick id
    uint32_t* slot_epoch;       // [n_buckets] call stamp of the touched set
    int4* coords;               // [n_bricks] block index + flags (bit 0: fresh)
    void* payload;              // [n_bricks][512] (sdf, w) pairs of T
    float* rgb;                 // [n_bricks][3][512] or null
    uint32_t* touched;          // [n_buckets] slots touched by the current call
    uint32_t* sb_bits;          // superblock occupancy, kSbSide^3 bits
    Ctrl* ctrl;
    uint32_t mask;              // n_buckets - 1
    uint32_t n_bricks;
    float vl, vbl, trunc, rmin, dmax, wmax;
    float rvbl;  // 1 / vbl rounded (div_index)
    int H, W;
    Camera cam;
};

// Brick payload: 512 (sdf, w) pairs, voxel v = x + 8y + 64z at pair v, so a
// raycast sample is one 8-byte (fp32) or 4-byte (fp16) load.
__host__ __device__ __forceinline__ void ld_pair(const float* b, int v, float& s, float& w)
{
    const float2 p = reinterpret_cast<const float2*>(b)[v];
    s = p.x;
    w = p.y;
}
__host__ __device__ __forceinline__ void ld_pair(const __half* b, int v, float& s, float& w)
{
    const __half2 p = reinterpret_cast<const __half2*>(b)[v];
    s = __low2float(p);
    w = __high2float(p);
}
__host__ __device__ __forceinline__ void st_pair(float* b, int v, float s, float w)
{
    reinterpret_cast<float2*>(b)[v] = make_float2(s, w);
}
__host__ __device__ __forceinline__ void st_pair(__half* b, int v, float s, float w)
{
    reinterpret_cast<__half2*>(b)[v] = __halves2half2(__float2half_rn(s), __float2half_rn(w));
}

// Linear probing from bucket b for `key`: brick id, or -1 when the key is not
// in the map (voxelMap2.find + valid_at).  Keys whose pool allocation failed
// carry an id >= n_bricks and count as absent.
__host__ __device__ __forceinline__ int probe_from(const GridView& g, uint32_t key, uint32_t b)
{
    for (uint32_t p = 0; p <= g.mask; ++p) {
        const unsigned long long s = g.slots[b];
        if (s == kSlotEmpty)
            return -1;
        if ((uint32_t)(s >> 32) == key) {
            const uint32_t id = (uint32_t)s;
            return id < g.n_bricks ? (int)id : -1;
        }
        b = (b + 1) & g.mask;
    }
    return -1;
}

__host__ __device__ __forceinline__ int find_brick(const GridView& g, int x, int y, int z)
{
    const uint32_t key = prime_hash(x, y, z);
    return probe_from(g, key, mix32(key) & g.mask);
}

// Probe result from the first slot word alone: the id, -1 (absent), or
// kProbeMore when the slot holds another key and probing must continue.
constexpr int kProbeMore = -2;
__host__ __device__ __forceinline__ int first_slot(const GridView& g, uint32_t key, unsigned long long s)
{
    const uint32_t id = (uint32_t)s;
    const int hit = id < g.n_bricks ? (int)id : -1;
    return s == kSlotEmpty ? -1 : ((uint32_t)(s >> 32) == key ? hit : kProbeMore);
}

__host__ __device__ __forceinline__ bool vin_ok(int x, int y, int z)
{
    return (unsigned)x < (unsigned)kVPE && (unsigned)y < (unsigned)kVPE && (unsigned)z < (unsigned)kVPE;
}

// Superblock occupancy.  A superblock is 8 x 8 x 8 blocks; the bitmap covers
// block indices [-kSbHalf * 8, kSbHalf * 8) per axis and a bit is set when
// any block inside was inserted.  A set bit may be stale (conservative).
constexpr int kSbSide = 256;
constexpr int kSbHalf = kSbSide / 2;
constexpr size_t kSbWords = (size_t)kSbSide * kSbSide * kSbSide / 32;

// Bit index of the superblock of block (x, y, z), or -1 outside the covered
// region.
__host__ __device__ __forceinline__ int sb_index(int x, int y, int z)
{
    const int sx = (x >> 3) + kSbHalf, sy = (y >> 3) + kSbHalf, sz = (z >> 3) + kSbHalf;
    const bool in = (unsigned)sx < (unsigned)kSbSide && (unsigned)sy < (unsigned)kSbSide &&
                    (unsigned)sz < (unsigned)kSbSide;
    return in ? sx + kSbSide * (sy + kSbSide * sz) : -1;
}

// The alias lattice of the reference key: the nonzero (dx, dy, dz) with
// 73856093 dx + 19349669 dy + 83492791 dz = 0 mod 2^32 and all |d| <= 2048
// are exactly these seven vectors and their negatives (found by solving for
// dz at every (dx, dy) in [-2048, 2048]^2; tests/test_alias_lattice.py).
__host__ __device__ __forceinline__ void alias_vector(int k, int& x, int& y, int& z)
{
    const int t[7][3] = {{152, 951, 211},     {1160, -1031, -67},  {1312, -80, 144},  {1464, 871, 355},
                         {1616, 1822, 566},   {304, 1902, 422},    {1008, -1982, -278}};
    x = t[k][0];
    y = t[k][1];
    z = t[k][2];
}

// True when no probe of a raycast whose probed blocks lie in the box
// [p_lo, p_hi] can find a key inserted by another block: no nonzero lattice
// vector lies in (inserted bounding box) - (probe box).  Then a probe at a
// block whose superblock bit is clear misses the hash, exactly as the
// reference's find would.
__host__ __device__ __forceinline__ bool alias_free(const int bmin[3], const int bmax[3], const int p_lo[3],
                                                    const int p_hi[3])
{
    if (bmin[0] > bmax[0])
        return true;  // nothing inserted
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = bmin[a] - p_hi[a];
        hi[a] = bmax[a] - p_lo[a];
        if (lo[a] < -2048 || hi[a] > 2048)
            return false;
    }
    for (int k = 0; k < 7; ++k) {
        int v[3];
        alias_vector(k, v[0], v[1], v[2]);
        bool pos = true, neg = true;
        for (int a = 0; a < 3; ++a) {
            pos = pos && lo[a] <= v[a] && v[a] <= hi[a];
            neg = neg && lo[a] <= -v[a] && -v[a] <= hi[a];
        }
        if (pos || neg)
            return false;
    }
    return true;
}

// Distinct voxels / keys a raycast reads (the compulsory-bytes terms of its
// roofline); touched only by the kStats instantiation of the march.
struct Footprint {
    uint32_t* vox_bits;          // [n_bricks][16] one bit per voxel
    unsigned long long* keys;    // open-addressing set of probed keys
    uint32_t key_mask;
    unsigned long long* counts;  // [0] distinct keys, [1] set overflow, [2] distinct voxels
};

#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline void fp_key(const Footprint& fp, uint32_t key)
{
    uint32_t b = mix32(key) & fp.key_mask;
    for (int p = 0; p < 256; ++p) {
        const unsigned long long s = fp.keys[b];
        if (s == (unsigned long long)key)
            return;
        if (s == kSlotEmpty) {
            const unsigned long long old = atomicCAS(&fp.keys[b], kSlotEmpty, (unsigned long long)key);
            if (old == kSlotEmpty) {
                atomicAdd(&fp.counts[0], 1ull);
                return;
            }
            if (old == (unsigned long long)key)
                return;
        }
        b = (b + 1) & fp.key_mask;
    }
    atomicAdd(&fp.counts[1], 1ull);
}
__device__ inline void fp_voxel(const Footprint& fp, int id, int v)
{
    atomicOr(&fp.vox_bits[(size_t)id * 16 + (v >> 5)], 1u << (v & 31));
}
#else
inline void fp_key(const Footprint&, uint32_t) {}
inline void fp_voxel(const Footprint&, int, int) {}
#endif

// Raycast constants of one call.
struct RayParams {
    float vl, vbl, dt, dmax;
    float rvl, rvbl;  // 1 / vl, 1 / vbl rounded (div_fast)
    float back;       // vbl - d_trunc, GPUVoxelGrid.cpp:613
    int H, W;
    int xcd_map;      // 1: each XCD takes one band of the image
    int use_sb;       // superblock shortcut allowed (alias certificate holds)
};

// getVBIndeces of a point: the fast path for all three axes, one shared
// fallback to the reference arithmetic.
__host__ __device__ __forceinline__ void vb3(f3 r, const RayParams& rp, int& x, int& y, int& z)
{
    bool a, b, c;
    x = div_fast(r.x, rp.rvbl, a);
    y = div_fast(r.y, rp.rvbl, b);
    z = div_fast(r.z, rp.rvbl, c);
    if (!(a && b && c)) {
        x = vb_index(r.x, rp.vbl);
        y = vb_index(r.y, rp.vbl);
        z = vb_index(r.z, rp.vbl);
    }
}

// getVinVBIndeces of a point in block (bx, by, bz), same scheme.
__host__ __device__ __forceinline__ void vin3(f3 r, int bx, int by, int bz, const RayParams& rp, int& x, int& y,
                                              int& z)
{
    bool a, b, c;
    x = div_fast(r.x - rp.vbl * (float)bx, rp.rvl, a);
    y = div_fast(r.y - rp.vbl * (float)by, rp.rvl, b);
    z = div_fast(r.z - rp.vbl * (float)bz, rp.rvl, c);
    if (!(a && b && c)) {
        x = vin_index(r.x, bx, rp.vbl, rp.vl);
        y = vin_index(r.y, by, rp.vbl, rp.vl);
        z = vin_index(r.z, bz, rp.vbl, rp.vl);
    }
}

struct RayResult {
    f3 p, c;
    int code;  // 1 zero crossing, 0 miss, -TSDF_RAY_UB_* reference UB
    unsigned int samples, probes;
};

constexpr int kRayUbVoxelIndex = 1;  // TSDF_RAY_UB_VOXEL_INDEX
constexpr int kRayUbMissingOut = 2;  // TSDF_RAY_UB_MISSING_OUT
constexpr int kSkipLook = 4;         // empty-block probes in flight per lane
constexpr int kInnerLook = 4;        // voxel reads in flight per lane

// Interior of superblock s in block units, inset so that any point whose
// approximate block coordinate q = p / vbl (error <= |q| 2^-22) lies within
// [lo, hi] has all three reference block indices inside the superblock.
__host__ __device__ __forceinline__ void set_sb_bounds(const int s[3], f3& lo, f3& hi)
{
    float* l = &lo.x;
    float* h = &hi.x;
    for (int a = 0; a < 3; ++a) {
        const float b0 = (float)(8 * s[a]), b1 = (float)(8 * s[a] + 8);
        const float e = fmaxf(fmaxf(fabsf(b0), fabsf(b1)), 1.0f) * 0x1p-17f;
        l[a] = b0 + e;
        h[a] = b1 - e;
    }
}

template <bool kStats>
__host__ __device__ __forceinline__ int probe_count(const GridView& g, const Footprint& fp, int x, int y, int z,
                                                    unsigned int& probes)
{
    ++probes;
    const uint32_t key = prime_hash(x, y, z);
    if (kStats)
        fp_key(fp, key);
    return probe_from(g, key, mix32(key) & g.mask);
}

__host__ __device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ f3 sel3(bool c, f3 a, f3 b) { return {c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }

// raycastingTag for pixel (i, j), GPUVoxelGrid.cpp:561-715.
//
// The loops are the reference's, with two changes of schedule that keep
// every probe, read and floating-point operation of the reference in its
// order:
//  - the empty-block skip (:596-604) computes its next kSkipLook positions
//    (which depend on no memory) and issues their first hash-slot loads at
//    once, then takes the steps in order until one hits;
//  - the voxel walk inside a block (:638-706) likewise loads its next
//    kInnerLook voxels at once and then applies the reference's decisions
//    to them in order.
// Steps are taken with selects rather than branches, so a wavefront stays
// on one instruction stream.  Reads that lie beyond the point where the
// reference stops are discarded.
template <typename T, bool kColor, bool kStats>
__host__ __device__ RayResult march(const GridView& g, const Pose& pose, const RayParams& rp, int i, int j,
                                    const Footprint& fp)
{
    RayResult out;
    out.p = {0.f, 0.f, 0.f};
    out.c = {0.f, 0.f, 0.f};
    out.code = 0;
    out.samples = 0;
    out.probes = 0;
    const float vl = rp.vl, vbl = rp.vbl, dt = rp.dt, dmax = rp.dmax;
    // createNormPcdTag, GPUVoxelGrid.cpp:717-722, and the ray set-up, :565-569
    f3 n = get_point(g.cam, i, j, 1.0f);
    const float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    n.x /= len;
    n.y /= len;
    n.z /= len;
    f3 rv = xform(n, pose.E);
    const f3 dir = {rv.x - pose.E[3], rv.y - pose.E[7], rv.z - pose.E[11]};
    float l = 0.1f;
    const float a0 = 1.0f - l;
    rv.x -= a0 * dir.x;
    rv.y -= a0 * dir.y;
    rv.z -= a0 * dir.z;
    const f3 skip = {vbl * dir.x, vbl * dir.y, vbl * dir.z};  // the reference's vbl * dir
    const f3 step = {vl * dir.x, vl * dir.y, vl * dir.z};     // and vl * dir

    bool Lp = false;
    bool sb_ok = false;  // sb_lo / sb_hi hold an empty superblock (block units, inset)
    f3 sb_lo = {0.f, 0.f, 0.f}, sb_hi = {0.f, 0.f, 0.f};
    while (l < dmax) {  // outer loop; done / Fn leave by return
        int bx, by, bz;
        vb3(rv, rp, bx, by, bz);
        int id = probe_count<kStats>(g, fp, bx, by, bz, out.probes);
        int count = 0;
        // ---- empty-block skip ----------------------------------------
        while (id < 0 && l < dmax) {
            // (a) steps into superblocks whose bit is clear: the probe
            // provably misses (alias_free), so only the reference's position
            // / length arithmetic is replayed; the bit is read once per
            // superblock entered
            if (rp.use_sb) {
                while (true) {
                    const f3 rn = add3(rv, skip);
                    const float qx = rn.x * rp.rvbl, qy = rn.y * rp.rvbl, qz = rn.z * rp.rvbl;
                    if (!(sb_ok && qx >= sb_lo.x && qx <= sb_hi.x && qy >= sb_lo.y && qy <= sb_hi.y &&
                          qz >= sb_lo.z && qz <= sb_hi.z)) {
                        int x, y, z;
                        vb3(rn, rp, x, y, z);
                        const int k = sb_index(x, y, z);
                        if (k < 0 || ((g.sb_bits[k >> 5] >> (k & 31)) & 1u))
                            break;  // not known empty: probe the hash
                        const int sbc[3] = {x >> 3, y >> 3, z >> 3};
                        set_sb_bounds(sbc, sb_lo, sb_hi);
                        sb_ok = true;
                    }
                    rv = rn;
                    l += vbl;
                    ++count;
                    ++out.probes;
                    if (!(l < dmax))
                        break;
                }
                if (!(l < dmax))
                    break;
            }
            // (b) kSkipLook further steps with their first hash slots loaded
            // at once, taken in order
            uint32_t key[kSkipLook], bk[kSkipLook];
            unsigned long long sl[kSkipLook];
            f3 r = rv;
#pragma unroll
            for (int k = 0; k < kSkipLook; ++k) {
                r = add3(r, skip);
                int x, y, z;
                vb3(r, rp, x, y, z);
                key[k] = prime_hash(x, y, z);
                bk[k] = mix32(key[k]) & g.mask;
                sl[k] = g.slots[bk[k]];
            }
            bool go = true, pend = false;
            uint32_t pkey = 0, pbk = 0;
#pragma unroll
            for (int k = 0; k < kSkipLook; ++k) {
                const int r0 = first_slot(g, key[k], sl[k]);
                if (kStats && go)
                    fp_key(fp, key[k]);
                const float ln = l + vbl;
                rv = sel3(go, add3(rv, skip), rv);
                l = go ? ln : l;
                count += go ? 1 : 0;
                out.probes += go ? 1u : 0u;
                id = go ? r0 : id;
                const bool more = go && r0 == kProbeMore;
                pend = pend || more;
                pkey = more ? key[k] : pkey;
                pbk = more ? bk[k] : pbk;
                go = go && r0 == -1 && ln < dmax;
            }
            if (pend)  // first slot held another key: finish the linear probe
                id = probe_from(g, pkey, (pbk + 1) & g.mask);
        }
        if (!(l < dmax))
            break;
        if (count > 0) {
            // back off and re-advance by d_trunc, GPUVoxelGrid.cpp:610-627
            const float back = rp.back;
            rv.x -= back * dir.x;
            rv.y -= back * dir.y;
            rv.z -= back * dir.z;
            l -= back;
            vb3(rv, rp, bx, by, bz);
            id = probe_count<kStats>(g, fp, bx, by, bz, out.probes);
            while (id < 0 && l < dmax) {
                rv.x += dt * dir.x;
                rv.y += dt * dir.y;
                rv.z += dt * dir.z;
                l += dt;
                vb3(rv, rp, bx, by, bz);
                id = probe_count<kStats>(g, fp, bx, by, bz, out.probes);
            }
        }
        if (!(l < dmax))
            break;
        // ---- voxel walk inside block (bx, by, bz) ----------------------
        const int cur = id;
        const T* brick = (const T*)g.payload + (size_t)cur * (2 * kVPB);
        int vx, vy, vz;
        vin3(rv, bx, by, bz, rp, vx, vy, vz);
        while (true) {
            // voxel index of positions rv, rv + vl dir, ...: -2 where the
            // loop test fails (range or block left), -1 where the index is
            // outside [0, 8) (reference UB), else x + 8y + 64z
            int vk[kInnerLook];
            float sk[kInnerLook], wk[kInnerLook];
            {
                f3 r = rv;
                float ll = l;
                bool same = l < dmax;
                int wx = vx, wy = vy, wz = vz;
#pragma unroll
                for (int k = 0; k < kInnerLook; ++k) {
                    if (k > 0) {
                        r = add3(r, step);
                        ll += vl;
                        int qx, qy, qz;
                        vb3(r, rp, qx, qy, qz);
                        same = same && qx == bx && qy == by && qz == bz && ll < dmax;
                        vin3(r, bx, by, bz, rp, wx, wy, wz);
                    }
                    vk[k] = !same ? -2 : (vin_ok(wx, wy, wz) ? wx + kVPE * wy + kVPE * kVPE * wz : -1);
                    ld_pair(brick, vk[k] >= 0 ? vk[k] : 0, sk[k], wk[k]);
                }
            }
            bool stop = false;
            int ending = 0;  // 1 left the block / range, 2 UB index, 3 Fn, 4 crossing
            float xs = 0.f, xw = 0.f;
            int xv = 0;
#pragma unroll
            for (int k = 0; k < kInnerLook; ++k) {
                const bool act = !stop;
                const bool proc = act && vk[k] >= 0;
                const bool wpos = wk[k] > 0, spos = sk[k] > 0;
                const bool neg = wpos && !spos;  // first voxel at or behind the surface
                const bool stp = proc && !neg;
                if (kStats && proc)
                    fp_voxel(fp, cur, vk[k]);
                out.samples += proc ? 1u : 0u;
                ending = (act && vk[k] == -2) ? 1 : ending;
                ending = (act && vk[k] == -1) ? 2 : ending;
                ending = (proc && neg) ? (Lp ? 4 : 3) : ending;
                xs = proc ? sk[k] : xs;
                xw = proc ? wk[k] : xw;
                xv = proc ? vk[k] : xv;
                Lp = stp ? wpos : Lp;
                rv = sel3(stp, add3(rv, step), rv);
                l = stp ? l + vl : l;
                stop = stop || !stp;
            }
            if (!stop) {  // walked past the window: position, block and index of rv
                int nx, ny, nz;
                vb3(rv, rp, nx, ny, nz);
                if (nx != bx || ny != by || nz != bz || !(l < dmax))
                    break;  // the inner loop test fails
                vin3(rv, bx, by, bz, rp, vx, vy, vz);
                continue;
            }
            if (ending == 1)
                break;  // back to the outer loop
            if (ending == 2) {
                out.code = -kRayUbVoxelIndex;
                return out;
            }
            if (ending == 3)
                return out;  // Fn: miss
            // zero crossing, GPUVoxelGrid.cpp:655-689
            const int ivx = xv & 7, ivy = (xv >> 3) & 7, ivz = xv >> 6;
            const f3 vm_in = {voxel_mid(ivx, bx, vbl, vl), voxel_mid(ivy, by, vbl, vl), voxel_mid(ivz, bz, vbl, vl)};
            rv.x -= vl * dir.x;
            rv.y -= vl * dir.y;
            rv.z -= vl * dir.z;
            l -= vl;
            int qx, qy, qz, wx, wy, wz;
            vb3(rv, rp, qx, qy, qz);
            vin3(rv, qx, qy, qz, rp, wx, wy, wz);
            const f3 vm_out = {voxel_mid(wx, qx, vbl, vl), voxel_mid(wy, qy, vbl, vl), voxel_mid(wz, qz, vbl, vl)};
            int ob = cur;
            if (qx != bx || qy != by || qz != bz) {
                ob = probe_count<kStats>(g, fp, qx, qy, qz, out.probes);
                if (ob < 0) {
                    out.code = -kRayUbMissingOut;
                    return out;
                }
            }
            if (!vin_ok(wx, wy, wz)) {
                out.code = -kRayUbVoxelIndex;
                return out;
            }
            const int vo_i = wx + kVPE * wy + kVPE * kVPE * wz;
            float os, ow;
            ld_pair((const T*)g.payload + (size_t)ob * (2 * kVPB), vo_i, os, ow);
            ++out.samples;
            if (kStats)
                fp_voxel(fp, ob, vo_i);
            const float xx = (ow * os / vl + xw * (1.0f - xs / vl)) / (xw + ow);
            f3 p;
            p.x = vm_out.x + xx * (vm_in.x - vm_out.x);
            p.y = vm_out.y + xx * (vm_in.y - vm_out.y);
            p.z = vm_out.z + xx * (vm_in.z - vm_out.z);
            const float sc = dot3(p, dir) - dot3(rv, dir);
            out.p.x = rv.x + sc * dir.x;
            out.p.y = rv.y + sc * dir.y;
            out.p.z = rv.z + sc * dir.z;
            if (kColor) {
                const float* co = g.rgb + (size_t)ob * (3 * kVPB);
                const float* cc = g.rgb + (size_t)cur * (3 * kVPB);
                out.c.x = (float)((0.5 * (double)co[vo_i] + 0.5 * (double)cc[xv]) / 255.0);
                out.c.y = (float)((0.5 * (double)co[vo_i + kVPB] + 0.5 * (double)cc[xv + kVPB]) / 255.0);
                out.c.z = (float)((0.5 * (double)co[vo_i + 2 * kVPB] + 0.5 * (double)cc[xv + 2 * kVPB]) / 255.0);
            }
            out.code = 1;
            return out;
        }
    }
    return out;
}

}  // namespace tsdf
