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
    int bbox_min[3];  // bounding box of every block ever inserted since reset
    int bbox_max[3];
    unsigned int sb_version;  // bumped when a superblock bit is set or the bits are cleared
    unsigned int pad_;
    unsigned int ray_q[2];    // deferred rays of a raycast, by raycast parity
};

// Statistics counters that many workgroups add to live outside Ctrl, each
// spread over kStatLanes cache lines: a wavefront or workgroup adds its count
// to one lane (its index mod kStatLanes) and the reader sums the lanes.
// Device-scope atomics on one address from every XCD queue up behind each
// other; spread over 32 lines they do not.
enum StatId : int {
    kStUpdated = 0,   // voxel updates (integrate, deintegrate, batch)
    kStValid,         // pixels with z > rmin (collection)
    kStNew,           // bricks allocated
    kStRayHits,       // raycast: zero crossings found
    kStRayUb,         // raycast: rays that ran into undefined behaviour
    kStRaySamples,    // raycast: voxel samples
    kStRayProbes,     // raycast: block lookups
    kStCount
};
constexpr int kStatLanes = 32;
constexpr int kStatStride = 16;  // unsigned long longs per lane: one 128-byte line
constexpr size_t kStatWords = (size_t)kStCount * kStatLanes * kStatStride;

// Empty-space distance over superblocks: for the superblocks inside the box
// of the set superblock bits, the Chebyshev distance (in superblocks, capped
// at kSbdfMax) to the nearest set bit.  Rebuilt by k_sbdf whenever
// Ctrl::sb_version moved; `version` is the Ctrl::sb_version it was built from.
constexpr int kSbdfMax = 31;
constexpr size_t kSbdfCap = (size_t)1 << 21;  // cells of the box
struct SbdfInfo {
    unsigned int version;
    int valid;     // 0: not usable (box outside the bitmap or larger than kSbdfCap)
    int lo[3];     // superblock coordinates of the box origin
    int dim[3];    // box size in superblocks (0: nothing inserted)
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
    uint8_t* sbdf;              // [kSbdfCap] empty-space distance over superblocks
    SbdfInfo* sbdf_info;
    Ctrl* ctrl;
    unsigned long long* stat;   // [kStCount][kStatLanes][kStatStride]
    uint32_t mask;              // n_buckets - 1
    uint32_t n_bricks;
    float vl, vbl, trunc, rmin, dmax, wmax;
    float rvbl;    // 1 / vbl rounded (render)
    double Rvbl;   // RN64(1 / vbl): exact block index through quot_rn
    int H, W;
    Camera cam;
    const float* colfac;  // [W] ((float)j - cj) / fx, the first factor of getPoint's x
    const float* rowfac;  // [H] ((float)i - ci) / fy, the first factor of getPoint's y
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

// Host builds (tests/cpp) may observe the loads of the kStats march: every
// loaded hash lookup and every voxel read, in march order.
struct HostTrace {
    void (*key)(void* ctx, uint32_t key);
    void (*voxel)(void* ctx, int id, int v);
    void* ctx;
};
inline HostTrace*& host_trace()
{
    static HostTrace* t = nullptr;
    return t;
}

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
inline void fp_key(const Footprint&, uint32_t key)
{
    if (HostTrace* t = host_trace())
        t->key(t->ctx, key);
}
inline void fp_voxel(const Footprint&, int id, int v)
{
    if (HostTrace* t = host_trace())
        t->voxel(t->ctx, id, v);
}
#endif

// Raycast constants of one call.
struct RayParams {
    float vl, vbl, dt, dmax;
    double Rvl, Rvbl;  // RN64(1 / vl), RN64(1 / vbl): exact indices through quot_rn
    float back;       // vbl - d_trunc, GPUVoxelGrid.cpp:613
    int H, W;
    int xcd_map;      // 1: each XCD takes one band of the image
    int order;        // tile order: 0 hardware (or xcd_map), 1 reversed, 2 `tile_order`
    const uint32_t* tile_order;  // [tiles] the order of the slowest-first schedule
    uint32_t* tile_cost;         // [tiles * waves] time of each wavefront, for the next schedule
    int use_sb;       // empty-space replay allowed (alias certificate and distance field hold)
    const uint8_t* df;  // the distance field and its box (filled in by the kernel)
    int df_lo[3], df_dim[3];
    int use_tb;       // tile depth bounds replay allowed (alias certificate holds)
    int budget;       // rounds of the per-lane march before a ray is deferred (0: never)
    unsigned long long* lc;  // the workgroup's LDS cache of lookup results, or null
};

// The LDS lookup cache of a raycast workgroup (a measurement build,
// TSDF_RAY_CACHE): kRayCache entries key << 32 | result, result = the brick
// id of the key or 0xFFFFFFFF where the key is not in the map.  Entry e starts
// as key e ^ 1, whose entry is not e, so it answers no lookup.  Entries are
// written with 64-bit LDS atomics: two keys of one entry never mix.
constexpr int kRayCache = 1024;
#ifndef TSDF_RAY_CACHE
#define TSDF_RAY_CACHE 0  // 1: k_raycast keeps the cache (the default build reads none)
#endif

// The first slot word of the lookup of `key`: from the cache where it holds
// the key (a word first_slot reads as the cached result), else from the
// table.
__host__ __device__ __forceinline__ unsigned long long lookup_slot(const GridView& g, const RayParams& rp,
                                                                 uint32_t key, uint32_t bk, bool& cached)
{
    cached = false;
#if defined(__HIP_DEVICE_COMPILE__) && TSDF_RAY_CACHE
    if (rp.lc) {
        const unsigned long long e = rp.lc[key & (kRayCache - 1)];
        if ((uint32_t)(e >> 32) == key) {
            cached = true;
            const uint32_t r = (uint32_t)e;
            return r == 0xFFFFFFFFu ? kSlotEmpty : (((unsigned long long)key << 32) | r);
        }
    }
#endif
    return g.slots[bk];
}

// Records the result of a finished lookup (id >= 0, or -1) in the cache.
__host__ __device__ __forceinline__ void cache_result(const RayParams& rp, uint32_t key, int id)
{
#if defined(__HIP_DEVICE_COMPILE__) && TSDF_RAY_CACHE
    if (rp.lc)
        atomicExch(&rp.lc[key & (kRayCache - 1)], ((unsigned long long)key << 32) | (uint32_t)id);
#else
    (void)rp;
    (void)key;
    (void)id;
#endif
}

// ---------------------------------------------------------------------------
// Tile depth bounds: the raycast's exact empty-space skip.
//
// Before a raycast, every allocated brick's box, grown by `grow` metres, is
// projected with the raycast's pose; each 8 x 8-pixel tile keeps the least
// camera depth of the grown boxes whose projection (grown by one pixel)
// touches it.  For a ray (i, j) of that tile and a march position whose ray
// parameter is l:
//  - the march's float position rv stays within `grow` of the exact point
//    P(l) = t + l dir (its float drift, bounded on the host from the number
//    of position updates a march can make and the float spacing at its
//    largest coordinate), so a brick whose box contains rv has a grown box
//    that contains P(l);
//  - P(l) projects to pixel (i, j) exactly and has camera depth l (dir . Ez)
//    = l n_z; a grown box in front of the camera containing P(l) projects
//    over (i, j), so it counted for the tile, with its least depth <= l n_z;
// hence l n_z below the tile's bound (less a float margin) means no brick
// holds rv, and the reference's lookup there misses (with the alias
// certificate: no other block's key equals rv's).  The march replays those
// steps without loading: the same positions, lengths and probe counts.
// Probe-run positions keep l >= 0.1 (the march starts there; a back-off
// returns less than the skip steps before it), so their depth is at least
// z_lo = 0.1 min(n_z) - grow: a brick is clipped there, and a brick that
// crosses that plane is bounded through its box in camera coordinates.  The
// map needs no clearing: an entry is (~epoch << 32) | bits(depth) under
// atomicMin, so this raycast's entries are below every older one and a tile
// without one reads as empty (every lookup of its rays misses).
// ---------------------------------------------------------------------------
#ifndef TSDF_TB_TILE
#define TSDF_TB_TILE 8
#endif
constexpr int kTbTile = TSDF_TB_TILE;  // tile side in pixels: the 8 x 8 pixels of a raycast wavefront

struct TileBounds {
    unsigned long long* z;  // [tiles_y * tiles_x]
    int tiles_x, tiles_y;
    uint32_t epoch;
    float grow;  // metres a brick box grows by
    float z_lo;  // least depth of a probe-run position
};

__host__ __device__ __forceinline__ unsigned long long tb_pack(uint32_t epoch, float z)
{
    union {
        float f;
        uint32_t u;
    } b;
    b.f = z;
    return ((unsigned long long)(~epoch) << 32) | b.u;
}

// The tile bound of this raycast (epoch), +inf for a tile no brick touched.
__host__ __device__ __forceinline__ float tb_depth(unsigned long long e, uint32_t epoch)
{
    union {
        float f;
        uint32_t u;
    } b;
    b.u = (uint32_t)e;
    return (uint32_t)(e >> 32) == ~epoch ? b.f : __builtin_inff();
}

// 1 / z to within a few ulps (the hardware reciprocal on the GPU); used only
// where a pixel of margin absorbs its error.
__host__ __device__ __forceinline__ float approx_rcp(float z)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(z);
#else
    return 1.0f / z;
#endif
}

// Bounds of brick (bx, by, bz): false when no probe-run position can lie in
// it (all of it nearer than z_lo, or outside the image), else its least
// depth beyond z_lo and the inclusive pixel range [r0, r1] x [c0, c1] its
// grown box may cover.
__host__ __device__ __forceinline__ bool brick_bounds(int bx, int by, int bz, float vbl, const Pose& P,
                                                      const Camera& cam, int H, int W, float grow, float z_lo,
                                                      float& zmin, int& r0, int& r1, int& c0, int& c1)
{
    const float lo[3] = {vbl * (float)bx - grow, vbl * (float)by - grow, vbl * (float)bz - grow};
    const float hi[3] = {vbl * (float)bx + vbl + grow, vbl * (float)by + vbl + grow, vbl * (float)bz + vbl + grow};
    float cx[8], cy[8], cz[8];
    float x0 = __builtin_inff(), x1 = -__builtin_inff(), y0 = __builtin_inff(), y1 = -__builtin_inff();
    float z0 = __builtin_inff(), z1 = -__builtin_inff();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float x = (k & 1) ? hi[0] : lo[0], y = (k & 2) ? hi[1] : lo[1], z = (k & 4) ? hi[2] : lo[2];
        cx[k] = x * P.Ei[0] + y * P.Ei[1] + z * P.Ei[2] + P.Ei[3];
        cy[k] = x * P.Ei[4] + y * P.Ei[5] + z * P.Ei[6] + P.Ei[7];
        cz[k] = x * P.Ei[8] + y * P.Ei[9] + z * P.Ei[10] + P.Ei[11];
        x0 = fminf(x0, cx[k]);
        x1 = fmaxf(x1, cx[k]);
        y0 = fminf(y0, cy[k]);
        y1 = fmaxf(y1, cy[k]);
        z0 = fminf(z0, cz[k]);
        z1 = fmaxf(z1, cz[k]);
    }
    if (!(z1 >= z_lo))
        return false;
    float rmin, rmax, cmin, cmax;
    if (z0 >= z_lo) {
        // in front of the clip plane: the projection is the convex hull of
        // the corners' projections
        rmin = cmin = __builtin_inff();
        rmax = cmax = -__builtin_inff();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            // a reciprocal within a few ulps: far inside the pixel of margin
            const float iz = approx_rcp(cz[k]);
            const float r = cy[k] * cam.fy * iz + cam.ci, c = cx[k] * cam.fx * iz + cam.cj;
            rmin = fminf(rmin, r);
            rmax = fmaxf(rmax, r);
            cmin = fminf(cmin, c);
            cmax = fmaxf(cmax, c);
        }
    } else {
        // crossing it: the part at depth [z_lo, z1] lies in the box
        // [x0, x1] x [y0, y1] x [z_lo, z1], over which x / z and y / z take
        // their extremes at the corners
        z0 = z_lo;
        const float i0 = approx_rcp(z0), i1 = approx_rcp(z1);
        const float u0 = fminf(x0 * i0, x0 * i1), u1 = fmaxf(x1 * i0, x1 * i1);
        const float v0 = fminf(y0 * i0, y0 * i1), v1 = fmaxf(y1 * i0, y1 * i1);
        cmin = fminf(u0 * cam.fx, u1 * cam.fx) + cam.cj;
        cmax = fmaxf(u0 * cam.fx, u1 * cam.fx) + cam.cj;
        rmin = fminf(v0 * cam.fy, v1 * cam.fy) + cam.ci;
        rmax = fmaxf(v0 * cam.fy, v1 * cam.fy) + cam.ci;
    }
    // one pixel of margin for the rounding of the projections
    rmin = fmaxf(rmin - 1.0f, -1.0f);
    cmin = fmaxf(cmin - 1.0f, -1.0f);
    rmax = fminf(rmax + 1.0f, (float)H);
    cmax = fminf(cmax + 1.0f, (float)W);
    if (!(rmin <= rmax) || !(cmin <= cmax))
        return false;  // off the image (or NaN)
    r0 = (int)ceilf(rmin);
    r1 = (int)floorf(rmax);
    c0 = (int)ceilf(cmin);
    c1 = (int)floorf(cmax);
    r0 = r0 < 0 ? 0 : r0;
    c0 = c0 < 0 ? 0 : c0;
    r1 = r1 > H - 1 ? H - 1 : r1;
    c1 = c1 > W - 1 ? W - 1 : c1;
    if (r0 > r1 || c0 > c1)
        return false;
    zmin = z0;
    return true;
}

// The ray parameter below which every lookup of ray (i, j) misses: the
// tile's bound over the ray's depth per unit length, less a float margin;
// d_max when no brick touched the tile.
__host__ __device__ __forceinline__ float tb_l_safe(float zt, float nz, float dmax)
{
    if (!(zt < __builtin_inff()))
        return dmax;
    const float l = (zt * (1.0f - 1e-5f) - 1e-5f) / nz;
    return l < dmax ? l : dmax;
}

// The reference's index rule on a quotient q = RN32(a / d) (exact through
// quot_rn): (int)q for a >= 0, (int)(q - 1) below.  A position on a ray lies
// within the pose translation plus d_max of the origin (|t| <= 1e5 m and
// d_max <= 1e3 m are enforced), so |q| stays far below 2^31 and the plain
// conversion truncates exactly like the reference's (int).
__host__ __device__ __forceinline__ int ray_index(float a, double R)
{
    const float q = quot_rn(a, R);
    return (int)((a >= 0) ? q : q - 1.0f);
}

// getVBIndeces of a point, GPUVoxelGrid.h:162-190.
__host__ __device__ __forceinline__ void vb3(f3 r, const RayParams& rp, int& x, int& y, int& z)
{
    x = ray_index(r.x, rp.Rvbl);
    y = ray_index(r.y, rp.Rvbl);
    z = ray_index(r.z, rp.Rvbl);
}

// getVinVBIndeces of a point in block (bx, by, bz), GPUVoxelGrid.h:192-221
// (p - vbl * vbInd in float, kokkosFunctions.h:737-740).
__host__ __device__ __forceinline__ void vin3(f3 r, int bx, int by, int bz, const RayParams& rp, int& x, int& y,
                                              int& z)
{
    x = ray_index(r.x - rp.vbl * (float)bx, rp.Rvl);
    y = ray_index(r.y - rp.vbl * (float)by, rp.Rvl);
    z = ray_index(r.z - rp.vbl * (float)bz, rp.Rvl);
}

struct RayResult {
    f3 p, c;
    int code;  // 1 zero crossing, 0 miss, -TSDF_RAY_UB_* reference UB, kRayDeferred
    unsigned int samples, probes;
};

// A march stopped at the top of its outer loop (code kRayDeferred) and
// resumed later by march_group: its position, length, the last sample's
// weight sign, the lookup loaded ahead for the next position and the counts
// so far.  Everything else of the march is recomputed from the pixel.
constexpr int kRayDeferred = 2;
struct RayState {
    float rx, ry, rz, l;
    uint32_t pix;          // i * W + j
    uint32_t flags;        // bit 0 Lp, bit 1 pre_valid
    uint32_t pre_key, pre_bk;
    unsigned long long pre_slot;
    uint32_t samples, probes;
    float zt;              // the tile depth bound of the ray (< 0: none)
    uint32_t pad[3];
};
static_assert(sizeof(RayState) == 64, "one 64-byte record per deferred ray");

constexpr int kRayUbVoxelIndex = 1;  // TSDF_RAY_UB_VOXEL_INDEX
constexpr int kRayUbMissingOut = 2;  // TSDF_RAY_UB_MISSING_OUT
#ifndef TSDF_SKIP_LOOK
#define TSDF_SKIP_LOOK 4
#endif
#ifndef TSDF_INNER_LOOK
#define TSDF_INNER_LOOK 4
#endif
constexpr int kSkipLook = TSDF_SKIP_LOOK;    // empty-block probes in flight per lane
constexpr int kInnerLook = TSDF_INNER_LOOK;  // voxel reads in flight per lane
static_assert(kInnerLook == 4, "the walk's block search is written for windows of 4 positions");

// Lower bound of the Chebyshev distance, in blocks, from block (x, y, z) to
// every inserted block: outside the box of set superblocks the distance to
// the box; inside it 8 (E - 1) + 1 for a distance of E >= 1 superblocks (the
// nearest inserted block is at least E - 1 whole superblocks and one block
// away), 1 in a set superblock.
__host__ __device__ __forceinline__ int df_box(const RayParams& rp, int x, int y, int z, int& idx)
{
    idx = 0;
    if (rp.df_dim[0] == 0)
        return 1 << 20;  // nothing inserted
    const int lx = rp.df_lo[0] * 8, ly = rp.df_lo[1] * 8, lz = rp.df_lo[2] * 8;
    const int hx = (rp.df_lo[0] + rp.df_dim[0]) * 8 - 1, hy = (rp.df_lo[1] + rp.df_dim[1]) * 8 - 1,
              hz = (rp.df_lo[2] + rp.df_dim[2]) * 8 - 1;
    int d = 0;
    d = lx - x > d ? lx - x : d;
    d = x - hx > d ? x - hx : d;
    d = ly - y > d ? ly - y : d;
    d = y - hy > d ? y - hy : d;
    d = lz - z > d ? lz - z : d;
    d = z - hz > d ? z - hz : d;
    if (d == 0)
        idx = ((x >> 3) - rp.df_lo[0]) + rp.df_dim[0] * (((y >> 3) - rp.df_lo[1]) + rp.df_dim[1] * ((z >> 3) - rp.df_lo[2]));
    return d;
}

__host__ __device__ __forceinline__ int df_dist(int dbox, int e)
{
    return dbox > 0 ? dbox : (e >= 1 ? 8 * (e - 1) + 1 : 1);
}

__host__ __device__ __forceinline__ int free_dist(const RayParams& rp, int x, int y, int z)
{
    int idx;
    const int d = df_box(rp, x, y, z, idx);
    return df_dist(d, d == 0 ? (int)rp.df[idx] : 0);
}

// Bit of superblock (sx, sy, sz) (superblock coordinates) in the bitmap.
__host__ __device__ __forceinline__ bool sb_bit(const uint32_t* bits, int sx, int sy, int sz)
{
    const int k = (sx + kSbHalf) + kSbSide * ((sy + kSbHalf) + kSbSide * (sz + kSbHalf));
    return ((bits[k >> 5] >> (k & 31)) & 1u) != 0;
}

// The distance field of the box [lo, lo + dim) of superblocks, capped at
// kSbdfMax, by the separable chessboard transform: along x the distance to
// the nearest set bit of the row, then D_y = min_dy max(|dy|, D_x), then
// D_z = min_dz max(|dz|, D_y) (windows of +-kSbdfMax).  Host build for the CPU
// harness; k_sbdf computes the same values on the GPU.  tmp: as large as out.
inline void sbdf_build_host(const uint32_t* sb_bits, const int lo[3], const int dim[3], uint8_t* out, uint8_t* tmp)
{
    const int nx = dim[0], ny = dim[1], nz = dim[2];
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            uint8_t* row = out + (size_t)nx * (y + (size_t)ny * z);
            int d = kSbdfMax;
            for (int x = 0; x < nx; ++x) {
                d = sb_bit(sb_bits, x + lo[0], y + lo[1], z + lo[2]) ? 0 : (d < kSbdfMax ? d + 1 : kSbdfMax);
                row[x] = (uint8_t)d;
            }
            d = kSbdfMax;
            for (int x = nx - 1; x >= 0; --x) {
                d = row[x] == 0 ? 0 : (d < kSbdfMax ? d + 1 : kSbdfMax);
                row[x] = (uint8_t)(row[x] < d ? row[x] : d);
            }
        }
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int best = kSbdfMax;
                for (int dy = -kSbdfMax; dy <= kSbdfMax; ++dy) {
                    const int yy = y + dy;
                    if (yy < 0 || yy >= ny)
                        continue;
                    const int a = dy < 0 ? -dy : dy;
                    const int v = out[x + (size_t)nx * (yy + (size_t)ny * z)];
                    const int m = a > v ? a : v;
                    best = m < best ? m : best;
                }
                tmp[x + (size_t)nx * (y + (size_t)ny * z)] = (uint8_t)best;
            }
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int best = kSbdfMax;
                for (int dz = -kSbdfMax; dz <= kSbdfMax; ++dz) {
                    const int zz = z + dz;
                    if (zz < 0 || zz >= nz)
                        continue;
                    const int a = dz < 0 ? -dz : dz;
                    const int v = tmp[x + (size_t)nx * (y + (size_t)ny * zz)];
                    const int m = a > v ? a : v;
                    best = m < best ? m : best;
                }
                out[x + (size_t)nx * (y + (size_t)ny * z)] = (uint8_t)best;
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
// One probing run of the march: the lookup at rv, then, while it misses and
// l < dmax, steps rv += delta, l += dl, each followed by a lookup -- the
// reference's loops at GPUVoxelGrid.cpp:589-604 (delta = vbl dir) and
// :610-627 (delta = d_trunc dir), with two changes of schedule that keep every
// lookup and floating-point operation in its order:
//  - kSkipLook positions (which depend on no memory) are computed and their
//    first hash slots loaded at once, then taken in order until one hits;
//  - with the empty-space field (rp.use_sb), a stretch of steps whose
//    lookups certainly miss is replayed without loading: every block within
//    Chebyshev distance d - 1 of the block of rv is absent, and k steps (each
//    at most vbl per axis) move the reference's block index by at most k + 3
//    per axis, so steps 1 .. d - 4 miss.
// Returns the brick id (>= 0) or -1; steps counts the steps taken, probes
// every lookup of the reference (replayed ones included).
//
// pre_valid: the first slot word of the lookup at rv (key pre_key, bucket
// pre_bk) was already loaded as pre_slot; a hit there costs no further load.
template <bool kStats>
__host__ __device__ __forceinline__ int probe_run(const GridView& g, const RayParams& rp, const Footprint& fp, f3& rv,
                                                  float& l, const f3 delta, float dl, int& steps, unsigned int& probes,
                                                  float l_safe, int& rounds, bool pre_valid = false,
                                                  uint32_t pre_key = 0, uint32_t pre_bk = 0,
                                                  unsigned long long pre_slot = 0)
{
    const float dmax = rp.dmax;
    int id = -1;
    bool first = true;  // the lookup at rv itself has not been made
    if (pre_valid) {
        const int r0 = first_slot(g, pre_key, pre_slot);
        if (kStats)
            fp_key(fp, pre_key);
        ++probes;
        id = r0 == kProbeMore ? probe_from(g, pre_key, (pre_bk + 1) & g.mask) : r0;
        cache_result(rp, pre_key, id);
        first = false;
        if (id >= 0)
            return id;
    }
    bool have = false;  // nrep holds the replay count for the block of rv
    int nrep = 0;
    while (first || (id < 0 && l < dmax)) {
        // tile depth bounds: the lookups at ray parameters below l_safe
        // miss (l_safe <= d_max, so every such miss is followed by a step)
        if (first && l < l_safe) {
            ++probes;
            first = false;
        }
        if (!first) {
            for (;;) {
                const float ln = l + dl;
                if (!(ln < l_safe))
                    break;
                rv = add3(rv, delta);
                l = ln;
                ++steps;
                ++probes;
            }
        }
        if (rp.use_sb && !first) {
            if (!have) {
                int cx, cy, cz;
                vb3(rv, rp, cx, cy, cz);
                nrep = free_dist(rp, cx, cy, cz) - 4;
            }
            for (int k = 0; k < nrep && l < dmax; ++k) {
                rv = add3(rv, delta);
                l += dl;
                ++steps;
                ++probes;
            }
            have = false;
            if (!(l < dmax))
                break;
        }
        // positions: rv itself first (on the first round), then steps
        uint32_t key[kSkipLook], bk[kSkipLook];
        unsigned long long sl[kSkipLook];
        bool cached[kSkipLook];
        int dbl = 0, ev = 0;
        f3 r = rv;
#pragma unroll
        for (int k = 0; k < kSkipLook; ++k) {
            if (!(first && k == 0))
                r = add3(r, delta);
            int x, y, z;
            vb3(r, rp, x, y, z);
            key[k] = prime_hash(x, y, z);
            bk[k] = mix32(key[k]) & g.mask;
            sl[k] = lookup_slot(g, rp, key[k], bk[k], cached[k]);
            if (k == kSkipLook - 1 && rp.use_sb) {
                int idx;
                dbl = df_box(rp, x, y, z, idx);
                ev = dbl == 0 ? (int)rp.df[idx] : 0;
            }
        }
        bool go = true, pend = false;
        ++rounds;
        uint32_t pkey = 0, pbk = 0;
#pragma unroll
        for (int k = 0; k < kSkipLook; ++k) {
            const int r0 = first_slot(g, key[k], sl[k]);
            if (!cached[k] && r0 != kProbeMore)
                cache_result(rp, key[k], r0);
            if (kStats && go)
                fp_key(fp, key[k]);
            const bool stp = go && !(first && k == 0);  // a step precedes this lookup
            const float ln = l + dl;
            rv = sel3(stp, add3(rv, delta), rv);
            l = stp ? ln : l;
            steps += stp ? 1 : 0;
            probes += go ? 1u : 0u;
            id = go ? r0 : id;
            const bool more = go && r0 == kProbeMore;
            pend = pend || more;
            pkey = more ? key[k] : pkey;
            pbk = more ? bk[k] : pbk;
            go = go && r0 == -1 && l < dmax;
        }
        first = false;
        if (rp.use_sb && go) {  // every lookup taken and missed: rv is the last position
            nrep = df_dist(dbl, ev) - 4;
            have = true;
        }
        if (pend) {  // first slot held another key: finish the linear probe
            id = probe_from(g, pkey, (pbk + 1) & g.mask);
            cache_result(rp, pkey, id);
        }
    }
    return id;
}

// The zero crossing, GPUVoxelGrid.cpp:655-689: rv is the first position at
// or behind the surface, in block (bx, by, bz) = brick cur, with sample
// (xs, xw) at voxel xv; out gets the point (and colour) or a UB code.
template <typename T, bool kColor, bool kStats>
__host__ __device__ __forceinline__ void crossing(const GridView& g, const RayParams& rp, const Footprint& fp, f3 rv,
                                                  float l, const f3 dir, int bx, int by, int bz, int cur, float xs,
                                                  float xw, int xv, RayResult& out)
{
    const float vl = rp.vl, vbl = rp.vbl;
    const int ivx = xv & 7, ivy = (xv >> 3) & 7, ivz = xv >> 6;
    const f3 vm_in = {voxel_mid(ivx, bx, vbl, vl), voxel_mid(ivy, by, vbl, vl), voxel_mid(ivz, bz, vbl, vl)};
    rv.x -= vl * dir.x;
    rv.y -= vl * dir.y;
    rv.z -= vl * dir.z;
    l -= vl;
    (void)l;
    int qx, qy, qz, wx, wy, wz;
    vb3(rv, rp, qx, qy, qz);
    vin3(rv, qx, qy, qz, rp, wx, wy, wz);
    const f3 vm_out = {voxel_mid(wx, qx, vbl, vl), voxel_mid(wy, qy, vbl, vl), voxel_mid(wz, qz, vbl, vl)};
    int ob = cur;
    if (qx != bx || qy != by || qz != bz) {
        ob = probe_count<kStats>(g, fp, qx, qy, qz, out.probes);
        if (ob < 0) {
            out.code = -kRayUbMissingOut;
            return;
        }
    }
    if (!vin_ok(wx, wy, wz)) {
        out.code = -kRayUbVoxelIndex;
        return;
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
}

// The ray of pixel (i, j), GPUVoxelGrid.cpp:565-569 and 717-722: the unit
// camera ray n, the start position rv (l = 0.1) and the direction dir.
__host__ __device__ __forceinline__ void ray_setup(const GridView& g, const Pose& pose, int i, int j, f3& n, f3& rv,
                                                   f3& dir)
{
    n = get_point(g.cam, i, j, 1.0f);
    const float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    n.x /= len;
    n.y /= len;
    n.z /= len;
    rv = xform(n, pose.E);
    dir = {rv.x - pose.E[3], rv.y - pose.E[7], rv.z - pose.E[11]};
    const float a0 = 1.0f - 0.1f;
    rv.x -= a0 * dir.x;
    rv.y -= a0 * dir.y;
    rv.z -= a0 * dir.z;
}

// budget > 0: after that many probe rounds and walk windows the march stops
// at the top of its outer loop, returns kRayDeferred and leaves its state in
// *st (march_group resumes it).
template <typename T, bool kColor, bool kStats>
__host__ __device__ RayResult march(const GridView& g, const Pose& pose, const RayParams& rp, int i, int j,
                                    const Footprint& fp, float zt = -1.0f, int budget = 0, RayState* st = nullptr)
{
    RayResult out;
    out.p = {0.f, 0.f, 0.f};
    out.c = {0.f, 0.f, 0.f};
    out.code = 0;
    out.samples = 0;
    out.probes = 0;
    const float vl = rp.vl, vbl = rp.vbl, dt = rp.dt, dmax = rp.dmax;
    // createNormPcdTag, GPUVoxelGrid.cpp:717-722, and the ray set-up, :565-569
    f3 n, rv, dir;
    ray_setup(g, pose, i, j, n, rv, dir);
    // the tile bound of this ray's tile (zt < 0: no replay)
    const float l_safe = (rp.use_tb && zt >= 0.0f) ? tb_l_safe(zt, n.z, dmax) : -1.0f;
    float l = 0.1f;
    int rounds = 0;
    const f3 skip = {vbl * dir.x, vbl * dir.y, vbl * dir.z};  // the reference's vbl * dir
    const f3 step = {vl * dir.x, vl * dir.y, vl * dir.z};     // and vl * dir

    bool Lp = false;
    // the lookup at rv, loaded ahead by the voxel walk that left a block
    bool pre_valid = false;
    uint32_t pre_key = 0, pre_bk = 0;
    unsigned long long pre_slot = 0;
    while (l < dmax) {  // outer loop; done / Fn leave by return
        if (budget > 0 && rounds >= budget && st) {
            st->rx = rv.x;
            st->ry = rv.y;
            st->rz = rv.z;
            st->l = l;
            st->pix = (uint32_t)(i * rp.W + j);
            st->flags = (Lp ? 1u : 0u) | (pre_valid ? 2u : 0u);
            st->pre_key = pre_key;
            st->pre_bk = pre_bk;
            st->pre_slot = pre_slot;
            st->samples = out.samples;
            st->probes = out.probes;
            st->zt = zt;
            out.code = kRayDeferred;
            return out;
        }
        // ---- the block lookup at rv and the empty-block skip ------------
        // (GPUVoxelGrid.cpp:589-604): probe rv, then steps of vbl * dir
        int count = 0;
        int id = probe_run<kStats>(g, rp, fp, rv, l, skip, vbl, count, out.probes, l_safe, rounds, pre_valid,
                                   pre_key, pre_bk, pre_slot);
        pre_valid = false;
        if (!(l < dmax))
            break;
        if (count > 0) {
            // back off and re-advance by d_trunc, GPUVoxelGrid.cpp:610-627:
            // probe the backed-off position, then steps of d_trunc * dir
            const float back = rp.back;
            rv.x -= back * dir.x;
            rv.y -= back * dir.y;
            rv.z -= back * dir.z;
            l -= back;
            const f3 dstep = {dt * dir.x, dt * dir.y, dt * dir.z};
            int steps = 0;
            id = probe_run<kStats>(g, rp, fp, rv, l, dstep, dt, steps, out.probes, l_safe, rounds);
        }
        if (!(l < dmax))
            break;
        int bx, by, bz;
        vb3(rv, rp, bx, by, bz);
        // ---- voxel walk inside block (bx, by, bz) ----------------------
        const int cur = id;
        const T* brick = (const T*)g.payload + (size_t)cur * (2 * kVPB);
        int vx, vy, vz;
        vin3(rv, bx, by, bz, rp, vx, vy, vz);
        while (true) {
            ++rounds;
            // positions r[0] = rv, r[k] = r[k - 1] + vl dir, and their
            // lengths, by the reference's additions
            f3 r[kInnerLook + 1];
            float ll[kInnerLook + 1];
            r[0] = rv;
            ll[0] = l;
#pragma unroll
            for (int k = 1; k <= kInnerLook; ++k) {
                r[k] = add3(r[k - 1], step);
                ll[k] = ll[k - 1] + vl;
            }
            // kl: the first of positions 1..kInnerLook outside block b
            // (kInnerLook + 1: none), and its block (e*).  Along each axis
            // the positions are monotone (float sums of a constant step) and
            // getVBIndeces is monotone (ray_index), and r[0] lies in b: once
            // a position is outside b every later one is, so the last
            // position decides and a binary search finds kl.
            int q4x, q4y, q4z;
            vb3(r[kInnerLook], rp, q4x, q4y, q4z);
            const bool in4 = q4x == bx && q4y == by && q4z == bz;
            int kl = kInnerLook + 1, ex = q4x, ey = q4y, ez = q4z;
            if (!in4) {
                int q2x, q2y, q2z;
                vb3(r[2], rp, q2x, q2y, q2z);
                const bool in2 = q2x == bx && q2y == by && q2z == bz;
                int qcx, qcy, qcz;
                vb3(sel3(in2, r[3], r[1]), rp, qcx, qcy, qcz);
                const bool inc = qcx == bx && qcy == by && qcz == bz;
                kl = in2 ? (inc ? 4 : 3) : (inc ? 2 : 1);
                const bool use2 = !in2 && inc;  // leaves at 2
                const bool usec = !inc;         // leaves at 1 or 3
                ex = usec ? qcx : (use2 ? q2x : q4x);
                ey = usec ? qcy : (use2 ? q2y : q4y);
                ez = usec ? qcz : (use2 ? q2z : q4z);
            }
            // the loop test of position k: in the block and l < d_max (the
            // lengths increase, so the first failure ends the window)
            bool e_found = false;  // the first position outside b is reached
#pragma unroll
            for (int k = 1; k <= kInnerLook; ++k)
                e_found = e_found || (kl == k && ll[k - 1] < dmax);
            // voxel index of positions 0..kInnerLook-1: -2 where the loop
            // test fails, -1 where the index is outside [0, 8) (reference
            // UB), else x + 8y + 64z
            int vk[kInnerLook];
            float sk[kInnerLook], wk[kInnerLook];
#pragma unroll
            for (int k = 0; k < kInnerLook; ++k) {
                int wx = vx, wy = vy, wz = vz;
                if (k > 0)
                    vin3(r[k], bx, by, bz, rp, wx, wy, wz);
                const bool same = k < kl && ll[k] < dmax;
                vk[k] = !same ? -2 : (vin_ok(wx, wy, wz) ? wx + kVPE * wy + kVPE * kVPE * wz : -1);
                ld_pair(brick, vk[k] >= 0 ? vk[k] : 0, sk[k], wk[k]);
            }
            uint32_t e_key = 0, e_bk = 0;
            unsigned long long e_slot = 0;
            if (e_found) {
                e_key = prime_hash(ex, ey, ez);
                e_bk = mix32(e_key) & g.mask;
                bool e_cached;
                e_slot = lookup_slot(g, rp, e_key, e_bk, e_cached);
            }
            bool stop = false;
            int ending = 0;  // 1 left the block / range, 2 UB index, 3 Fn, 4 crossing
            int nst = 0;     // positions taken (the march goes on from r[nst])
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
                nst = stp ? k + 1 : nst;
                stop = stop || !stp;
            }
#pragma unroll
            for (int k = 1; k <= kInnerLook; ++k) {
                rv = sel3(nst == k, r[k], rv);
                l = nst == k ? ll[k] : l;
            }
            if (!stop) {  // walked past the window: rv = r[kInnerLook]
                if (!in4 || !(l < dmax)) {
                    // the inner loop test fails; rv is the window's next
                    // position, whose lookup was loaded if it left the block
                    pre_valid = e_found;
                    pre_key = e_key;
                    pre_bk = e_bk;
                    pre_slot = e_slot;
                    break;
                }
                vin3(rv, bx, by, bz, rp, vx, vy, vz);
                continue;
            }
            if (ending == 1) {  // back to the outer loop from the first position that left
                pre_valid = e_found;
                pre_key = e_key;
                pre_bk = e_bk;
                pre_slot = e_slot;
                break;
            }
            if (ending == 2) {
                out.code = -kRayUbVoxelIndex;
                return out;
            }
            if (ending == 3)
                return out;  // Fn: miss
            // zero crossing, GPUVoxelGrid.cpp:655-689
            crossing<T, kColor, kStats>(g, rp, fp, rv, l, dir, bx, by, bz, cur, xs, xw, xv, out);
            return out;
        }
    }
    return out;
}

}  // namespace tsdf
