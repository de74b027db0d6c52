// tsdf_hip.hip -- MI355X (gfx950) voxel-hashed TSDF engine behind the C ABI of
// include/tsdf.h.
//
// Device data layout (all in HBM, owned by the handle):
//   slots      [n_buckets] u64   open-addressing hash, (key << 32) | brick id
//   slot_epoch [n_buckets] u32   call stamp; replaces the reference's
//                                alreadyIntegrated flag + O(capacity) reset
//                                (GPUVoxelGrid.cpp:76-78, 552-559)
//   coords     [n_bricks] int4   block index (x, y, z) + flags (bit 0: fresh)
//   payload    [n_bricks][2][512] T   per brick: sdf[512] then w[512]
//                                (T = float or half; voxel v = x + 8y + 64z)
//   rgb        [n_bricks][3][512] float   optional colour, channel-planar
//   touched    [n_buckets] u32   slots claimed by the current call
//   ptmap      [H][W] float4     world point of every pixel of the current
//                                frame; x = NaN where camera z <= rmin
//
// Kernels per integrate call: k_collect (allocation, one pass over pixels)
// and k_integrate (one wavefront per touched brick).  The brick count never
// leaves the device: k_integrate is a persistent grid that reads it.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tsdf.h"
#include "tsdf_math.hpp"
#include "tsdf_march.hpp"
#include "tsdf_march_group.hpp"

#pragma clang fp contract(off)

using namespace tsdf;

namespace {

// ---------------------------------------------------------------------------
// device-side state
// ---------------------------------------------------------------------------



#ifndef TSDF_DEPTH_GATHER
#define TSDF_DEPTH_GATHER 0  // plain update: 1 gathers the depth and rebuilds the point, 0 gathers the point map
#endif
#ifndef TSDF_DEPTH_TILE
#define TSDF_DEPTH_TILE 0  // plain update: 1 stages each brick's point-map rectangle in LDS
#endif

// One integrate (+1) or deintegrate (-1) operation.
struct OpDesc {
    const uint16_t* depth;
    const uint8_t* rgb;
    float4* ptmap;
    int pitch;
    int sign;
    Pose pose;
};

// (v + 0.5) * vl in double for v = 0..7 (the reference's promotion at
// GPUVoxelGrid.cpp:521).
struct Offsets {
    double v[8];
};

constexpr int kThreads = 256;
#ifndef TSDF_TILE_W
#define TSDF_TILE_W 64
#endif
#ifndef TSDF_TILE_H
#define TSDF_TILE_H 8
#endif
#ifndef TSDF_LDS_SET
#define TSDF_LDS_SET 1024
#endif
// Collection tile: 64 x 8 pixels = 8 patches of 8 x 8, two per wavefront.
// Its block set holds kLdsSet blocks (8 KiB); a tile sees several tens of
// blocks on average, but grazing views put hundreds in one tile, and a full
// set falls back to claiming blocks one by one, which costs far more than
// the set.
constexpr int kTileW = TSDF_TILE_W;
constexpr int kTileH = TSDF_TILE_H;
constexpr int kLdsSet = TSDF_LDS_SET;
// A batch pass collects several ops at once (grid.z = op), which fills the
// GPU with larger tiles; larger tiles also see a block fewer times per op.
#ifndef TSDF_BATCH_TILE_W
#define TSDF_BATCH_TILE_W 64
#endif
#ifndef TSDF_BATCH_TILE_H
#define TSDF_BATCH_TILE_H 16
#endif
constexpr int kBatchTileW = TSDF_BATCH_TILE_W;
constexpr int kBatchTileH = TSDF_BATCH_TILE_H;
constexpr int kBatchLdsSet = 2048;  // four times the plain tile's pixels
constexpr int kMaxOps = 256;    // ops per batched pass
constexpr int kMaskWords = kMaxOps / 32;

// What the host stages for one batch call (pinned copy -> device copy): the
// first call stamp of the batch and its op descriptors.
struct BatchStage {
    uint32_t epoch_base;
    uint32_t pad[3];
    OpDesc ops[kMaxOps];
};
constexpr int kMaxAxisBlocks = 64;
constexpr float kF16MaxWeight = 2048.0f;  // largest weight fp16 counts exactly

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------


__device__ __forceinline__ unsigned long long pack_coord(int x, int y, int z)
{
    return ((unsigned long long)(x & 0x1FFFFF) << 42) | ((unsigned long long)(y & 0x1FFFFF) << 21) |
           (unsigned long long)(z & 0x1FFFFF);
}
__device__ __forceinline__ int unpack21(unsigned long long v)
{
    int r = (int)(v & 0x1FFFFF);
    return (r ^ 0x100000) - 0x100000;
}
__device__ __forceinline__ bool packable(int v) { return v >= -(1 << 20) && v < (1 << 20); }

// Appends `val` for every lane with `need` using one atomic per wavefront.
__device__ __forceinline__ void wave_append(bool need, uint32_t val, unsigned int* counter,
                                            uint32_t* list)
{
    const unsigned long long m = __ballot(need);
    if (m == 0)
        return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if (lane == leader)
        base = atomicAdd(counter, (unsigned int)__popcll(m));
    base = __shfl(base, leader);
    if (need)
        list[base + __popcll(m & ((1ull << lane) - 1ull))] = val;
}

__device__ __forceinline__ unsigned int wave_sum(unsigned int v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ unsigned int workgroup_index()
{
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// Lane `who` (mod kStatLanes) of statistics counter `id`.
__device__ __forceinline__ unsigned long long* stat_lane(const GridView& g, int id, unsigned int who)
{
    return g.stat + ((size_t)id * kStatLanes + who % kStatLanes) * kStatStride;
}

// Widens the inserted-block box to (x, y, z).  The box only grows, so a
// coordinate inside a value read earlier is inside the current box too and
// needs no atomic (most inserts fall inside it).
__device__ __forceinline__ void grow_bbox(Ctrl* c, int x, int y, int z)
{
    const int v[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (v[a] < __hip_atomic_load(&c->bbox_min[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&c->bbox_min[a], v[a]);
        if (v[a] > __hip_atomic_load(&c->bbox_max[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&c->bbox_max[a], v[a]);
    }
}

// Finds `key` or inserts it.  The thread whose CAS claims an empty slot
// takes the next brick id and publishes it; concurrent finders of the same
// key only need the slot index.  Returns the slot or -1 if the table is full.
__device__ int find_or_insert(const GridView& g, uint32_t key, int x, int y, int z)
{
    uint32_t b = mix32(key) & g.mask;
    const unsigned long long keyhi = (unsigned long long)key << 32;
    for (uint32_t p = 0; p <= g.mask; ++p) {
        unsigned long long s = __hip_atomic_load(&g.slots[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s == kSlotEmpty) {
            const unsigned long long old = atomicCAS(&g.slots[b], kSlotEmpty, keyhi | kIdPending);
            if (old == kSlotEmpty) {
                const unsigned int id = atomicAdd(&g.ctrl->alloc_top, 1u);
                unsigned int idv = kIdOverflow;
                if (id < g.n_bricks) {
                    g.coords[id] = make_int4(x, y, z, 1);
                    idv = id;
                    atomicAdd(stat_lane(g, kStNew, workgroup_index()), 1ull);
                }
                {
                    // every inserted key counts for the raycast's empty-space
                    // certificate, whether or not its brick was allocated
                    const int sb = sb_index(x, y, z);
                    if (sb >= 0 && !(atomicOr(&g.sb_bits[sb >> 5], 1u << (sb & 31)) & (1u << (sb & 31))))
                        atomicAdd(&g.ctrl->sb_version, 1u);  // a new superblock: the distance field is stale
                    grow_bbox(g.ctrl, x, y, z);
                }
                if (id >= g.n_bricks)
                    atomicAdd(&g.ctrl->overflow, 1u);
                atomicExch(&g.slots[b], keyhi | idv);
                return (int)b;
            }
            s = old;
        }
        if ((uint32_t)(s >> 32) == key)
            return (int)b;
        b = (b + 1) & g.mask;
    }
    return -1;
}

// Marks the block as touched by this call (and by op `op` in a batch).
// Returns true for the one caller that must append the slot to the list.
template <bool kBatch>
__device__ __forceinline__ bool claim(const GridView& g, int x, int y, int z, uint32_t epoch,
                                      uint32_t* op_mask, int op, uint32_t* slot_out)
{
    // The key's home bucket and its stamp are read together: at the table's
    // load factor the key is usually there, and then a block that another
    // tile already claimed costs one round trip.
    const uint32_t key = prime_hash(x, y, z);
    const uint32_t b0 = mix32(key) & g.mask;
    const unsigned long long s0 = __hip_atomic_load(&g.slots[b0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t e0 = __hip_atomic_load(&g.slot_epoch[b0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // a batch op's bit, read with the home bucket: most (block, op) pairs are
    // claimed by several tiles, and only the first needs the atomic
    const uint32_t bit = 1u << (op & 31);
    const uint32_t m0 =
        kBatch ? __hip_atomic_load(&op_mask[(size_t)b0 * kMaskWords + (op >> 5)], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT)
               : 0u;
    int s;
    uint32_t pre, mword;
    if (s0 != kSlotEmpty && (uint32_t)(s0 >> 32) == key) {
        s = (int)b0;
        pre = e0;
        mword = m0;
    } else {
        s = find_or_insert(g, key, x, y, z);
        if (s < 0) {
            atomicAdd(&g.ctrl->overflow, 1u);
            return false;
        }
        pre = __hip_atomic_load(&g.slot_epoch[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mword = 0u;
    }
    // bits are only set during collection, so one seen set stays set
    if (kBatch && !(mword & bit))
        atomicOr(&op_mask[(size_t)s * kMaskWords + (op >> 5)], bit);
    *slot_out = (uint32_t)s;
    if (pre == epoch)
        return false;  // another tile of this call claimed it already
    return atomicExch(&g.slot_epoch[s], epoch) != epoch;
}

// Bits [a, b) of one 64-bit word, 0 <= a < b <= 64.
__device__ __forceinline__ unsigned long long bit_range(int a, int b)
{
    const unsigned long long hi = b >= 64 ? ~0ull : ((1ull << b) - 1ull);
    return hi & ~((1ull << a) - 1ull);
}

// Sets bits [base, base + n) of the 128-bit value (lo, hi), base + n <= 128.
__device__ __forceinline__ void set_bits128(unsigned long long& lo, unsigned long long& hi, int base, int n)
{
    const int e = base + n;
    if (base < 64)
        lo |= bit_range(base, e < 64 ? e : 64);
    if (e > 64)
        hi |= bit_range(base > 64 ? base - 64 : 0, e - 64);
}

template <int kSet>
__device__ __forceinline__ bool lds_insert(unsigned long long* set, unsigned long long v)
{
    uint32_t h = mix32((uint32_t)v ^ ((uint32_t)(v >> 32) * 0x9E3779B1u)) & (kSet - 1);
    for (int p = 0; p < 32; ++p) {
        const unsigned long long cur = set[h];  // most candidates are already present
        if (cur == v)
            return true;
        if (cur != kSlotEmpty) {
            h = (h + 1) & (kSet - 1);
            continue;
        }
        const unsigned long long old = atomicCAS(&set[h], kSlotEmpty, v);
        if (old == kSlotEmpty || old == v)
            return true;
        h = (h + 1) & (kSet - 1);
    }
    return false;
}

// ---------------------------------------------------------------------------
// k_reset: empty table, zero counters
// ---------------------------------------------------------------------------

__global__ void k_reset(GridView g, uint32_t* op_mask, size_t op_mask_words)
{
    const size_t n = (size_t)g.mask + 1;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        g.slots[i] = kSlotEmpty;
        g.slot_epoch[i] = 0u;
    }
    if (op_mask)
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < op_mask_words; i += stride)
            op_mask[i] = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < kSbWords; i += stride)
        g.sb_bits[i] = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < kStatWords; i += stride)
        g.stat[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Ctrl z;
        memset(&z, 0, sizeof(z));
        for (int a = 0; a < 3; ++a) {
            z.bbox_min[a] = 0x7FFFFFFF;
            z.bbox_max[a] = (int)0x80000000u;
        }
        z.sb_version = g.ctrl->sb_version + 1u;  // the bits were cleared
        *g.ctrl = z;
    }
}

// ---------------------------------------------------------------------------
// k_collect: block allocation (vbCollectionTag, GPUVoxelGrid.cpp:471-514)
// ---------------------------------------------------------------------------
// One workgroup per kTileW x kTileH pixel tile.  Each valid pixel enumerates the
// blocks of its +-trunc box into an LDS set, so the global table sees each
// block once per tile; then the tile's set is claimed against the global
// table with wavefront-aggregated appends to the touched list.  The pixel's
// world point goes to ptmap for k_integrate (the reference's inputPcd,
// GPUVoxelGrid.h:303-307, transformed as at GPUVoxelGrid.cpp:356).
template <bool kBatch, int kTW, int kTH, int kSet>
__device__ void collect_tile(const GridView& g, const OpDesc& op, uint32_t epoch, int par,
                             uint32_t* op_mask, int op_index)
{
    __shared__ unsigned long long lset[kSet];
    __shared__ unsigned int s_valid;
    for (int e = threadIdx.x; e < kSet; e += kThreads)
        lset[e] = kSlotEmpty;
    if (threadIdx.x == 0)
        s_valid = 0;
    __syncthreads();

    // Each wavefront takes 8 x 8 pixel patches of the tile (lane =
    // pixel of the patch).  The blocks of a patch are the union of its
    // pixels' +-trunc boxes; when their common bounding box has at most 128
    // cells, that union is built exactly as a 128-bit mask over the box (each
    // lane ORs in its own box, then the wave ORs the masks) and every lane
    // inserts at most two distinct blocks into the tile's set.  Larger unions
    // fall back to per-pixel enumeration.
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    unsigned int nvalid = 0;
    for (int pt_idx = wv; pt_idx < (kTW / 8) * (kTH / 8); pt_idx += kThreads / 64) {
        const int i = blockIdx.y * kTH + (pt_idx / (kTW / 8)) * 8 + (lane >> 3);
        const int j = blockIdx.x * kTW + (pt_idx % (kTW / 8)) * 8 + (lane & 7);
        const bool in = i < g.H && j < g.W;
        int x0 = 0x7FFFFFFF, y0 = 0x7FFFFFFF, z0 = 0x7FFFFFFF;
        int x1 = (int)0x80000000u, y1 = (int)0x80000000u, z1 = (int)0x80000000u;
        bool box = false;  // this pixel contributes a box
        if (in) {
            const float d = (float)op.depth[(size_t)i * op.pitch + j] / 1000.0f;
            const f3 p = get_point(g.cam, i, j, d);
            // pixels without a measurement (z <= rmin) carry x = NaN: any sdf
            // computed from them is NaN and fails the |sdf| < trunc test
            float4 o = make_float4(__builtin_nanf(""), 0.f, 0.f, 0.f);
            if (p.z > g.rmin) {
                ++nvalid;
                const f3 pw = xform(p, op.pose.E);
                o = make_float4(pw.x, pw.y, pw.z, 1.0f);
                const int bx0 = vb_index_q(pw.x - g.trunc, g.Rvbl), bx1 = vb_index_q(pw.x + g.trunc, g.Rvbl);
                const int by0 = vb_index_q(pw.y - g.trunc, g.Rvbl), by1 = vb_index_q(pw.y + g.trunc, g.Rvbl);
                const int bz0 = vb_index_q(pw.z - g.trunc, g.Rvbl), bz1 = vb_index_q(pw.z + g.trunc, g.Rvbl);
                const bool sane = (long long)bx1 - bx0 < kMaxAxisBlocks && (long long)by1 - by0 < kMaxAxisBlocks &&
                                  (long long)bz1 - bz0 < kMaxAxisBlocks;
                const bool pk = packable(bx0) && packable(bx1) && packable(by0) && packable(by1) && packable(bz0) &&
                                packable(bz1);
                if (sane && pk) {
                    x0 = bx0; x1 = bx1; y0 = by0; y1 = by1; z0 = bz0; z1 = bz1;
                    box = true;
                } else if (!sane) {
                    atomicAdd(&g.ctrl->overflow, 1u);
                } else {
                    // a block index beyond the packable range: claim directly
                    for (int x = bx0; x <= bx1; ++x)
                        for (int y = by0; y <= by1; ++y)
                            for (int z = bz0; z <= bz1; ++z) {
                                uint32_t slot;
                                if (claim<kBatch>(g, x, y, z, epoch, op_mask, op_index, &slot))
                                    g.touched[atomicAdd(&g.ctrl->touched_count[par], 1u)] = slot;
                            }
                }
            }
            if (kBatch || !TSDF_DEPTH_GATHER)
                op.ptmap[(size_t)i * g.W + j] = o;
        }
        if (__ballot(box) == 0ull)
            continue;
        // bounding box of the patch's boxes (wave reduction)
        int ux0 = x0, uy0 = y0, uz0 = z0, ux1 = x1, uy1 = y1, uz1 = z1;
        for (int m = 32; m > 0; m >>= 1) {
            ux0 = min(ux0, __shfl_xor(ux0, m));
            uy0 = min(uy0, __shfl_xor(uy0, m));
            uz0 = min(uz0, __shfl_xor(uz0, m));
            ux1 = max(ux1, __shfl_xor(ux1, m));
            uy1 = max(uy1, __shfl_xor(uy1, m));
            uz1 = max(uz1, __shfl_xor(uz1, m));
        }
        const long long ex = (long long)ux1 - ux0 + 1, ey = (long long)uy1 - uy0 + 1, ez = (long long)uz1 - uz0 + 1;
        if (ex * ey * ez <= 128) {
            const int nx = (int)ex, nxy = (int)(ex * ey);
            unsigned long long mlo = 0ull, mhi = 0ull;
            if (box)
                for (int z = z0; z <= z1; ++z)
                    for (int y = y0; y <= y1; ++y)
                        set_bits128(mlo, mhi, (x0 - ux0) + nx * (y - uy0) + nxy * (z - uz0), x1 - x0 + 1);
            for (int m = 32; m > 0; m >>= 1) {
                mlo |= __shfl_xor(mlo, m);
                mhi |= __shfl_xor(mhi, m);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = lane + 64 * h;
                const unsigned long long word = h ? mhi : mlo;
                if (!((word >> lane) & 1ull))
                    continue;
                const int cz = c / nxy, rem = c - cz * nxy;
                const int cy = rem / nx, cx = rem - cy * nx;
                const int x = ux0 + cx, y = uy0 + cy, z = uz0 + cz;
                if (!lds_insert<kSet>(lset, pack_coord(x, y, z))) {  // set full: claim directly
                    uint32_t slot;
                    if (claim<kBatch>(g, x, y, z, epoch, op_mask, op_index, &slot))
                        g.touched[atomicAdd(&g.ctrl->touched_count[par], 1u)] = slot;
                }
            }
        } else if (box) {
            for (int x = x0; x <= x1; ++x)
                for (int y = y0; y <= y1; ++y)
                    for (int z = z0; z <= z1; ++z) {
                        if (!lds_insert<kSet>(lset, pack_coord(x, y, z))) {  // set full: claim directly
                            uint32_t slot;
                            if (claim<kBatch>(g, x, y, z, epoch, op_mask, op_index, &slot))
                                g.touched[atomicAdd(&g.ctrl->touched_count[par], 1u)] = slot;
                        }
                    }
        }
    }
    nvalid = wave_sum(nvalid);
    if (__lane_id() == 0 && nvalid)
        atomicAdd(&s_valid, nvalid);
    __syncthreads();
    if (threadIdx.x == 0 && s_valid)
        atomicAdd(stat_lane(g, kStValid, workgroup_index()), (unsigned long long)s_valid);

    // compact the tile's distinct blocks, then claim them all in one round
    // (one dependent chain of global accesses per tile rather than one per
    // stripe of the set)
    __shared__ unsigned long long uniq[kSet];
    __shared__ unsigned int n_uniq;
    if (threadIdx.x == 0)
        n_uniq = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < kSet; e += kThreads) {
        const unsigned long long v = lset[e];
        const bool has = v != kSlotEmpty;
        const unsigned long long m = __ballot(has);
        unsigned int base = 0;
        if (__lane_id() == 0 && m)
            base = atomicAdd(&n_uniq, (unsigned int)__popcll(m));
        base = __shfl(base, 0);
        if (has)
            uniq[base + __popcll(m & ((1ull << __lane_id()) - 1ull))] = v;
    }
    __syncthreads();
    // The slots this tile must append are gathered in LDS (lset is free
    // now) and appended with one global atomic for the whole tile: the
    // touched-list counter is one address that every tile adds to.
    const unsigned int nu = n_uniq;
    uint32_t* app = reinterpret_cast<uint32_t*>(lset);  // 2 * kSet >= nu entries
    __shared__ unsigned int n_app, app_base;
    if (threadIdx.x == 0)
        n_app = 0;
    __syncthreads();
    for (unsigned int e = threadIdx.x; e < ((nu + 63u) & ~63u); e += kThreads) {
        bool need = false;
        uint32_t slot = 0;
        if (e < nu) {
            const unsigned long long v = uniq[e];
            need = claim<kBatch>(g, unpack21(v >> 42), unpack21(v >> 21), unpack21(v), epoch, op_mask, op_index,
                                 &slot);
        }
        wave_append(need, slot, &n_app, app);
    }
    __syncthreads();
    const unsigned int na = n_app;
    if (threadIdx.x == 0 && na)
        app_base = atomicAdd(&g.ctrl->touched_count[par], na);
    __syncthreads();
    for (unsigned int e = threadIdx.x; e < na; e += kThreads)
        g.touched[app_base + e] = app[e];
}

__global__ __launch_bounds__(kThreads) void k_collect(GridView g, OpDesc op, uint32_t epoch, int par)
{
    collect_tile<false, kTileW, kTileH, kLdsSet>(g, op, epoch, par, nullptr, 0);
}

// A batch pass reads its call stamp from device memory (*epoch_base + pass)
// so that a captured graph of the whole batch can be replayed with new
// stamps.
__global__ __launch_bounds__(kThreads) void k_collect_batch(GridView g, const OpDesc* ops,
                                                            const uint32_t* epoch_base, int pass, uint32_t* op_mask)
{
    const uint32_t epoch = *epoch_base + (uint32_t)pass;
    collect_tile<true, kBatchTileW, kBatchTileH, kBatchLdsSet>(g, ops[blockIdx.z], epoch, (int)(epoch & 1u), op_mask,
                                                               blockIdx.z);
}

// ---------------------------------------------------------------------------
// per-voxel update (updateSingleVoxel[Color] / deUpdateSingleVoxel[Color],
// GPUVoxelGrid.cpp:345-469) on register-resident values
// ---------------------------------------------------------------------------

// Row-major pixel of the voxel centre with camera coordinates c, or -1 when
// it falls outside the image (GPUVoxelGrid.cpp:347-352).  The reciprocal
// comes from v_rcp_f32 and is certified before use (rcp_certified), so the
// result rests on IEEE fma / mul / add / floor only, not on the accuracy of
// the hardware reciprocal.
__device__ __forceinline__ int project_voxel(const GridView& g, float cx, float cy, float cz)
{
    const float r = __builtin_amdgcn_rcpf(cz);
    const float ny = cy * g.cam.fy, nx = cx * g.cam.fx;
    int i, j;
    // both axes are evaluated (no short circuit), then one branch
    const bool ok = ((int)rcp_certified(r, cz) & (int)pix_axis_fast(ny, r, g.cam.ci, i) &
                     (int)pix_axis_fast(nx, r, g.cam.cj, j)) != 0;
    if (!ok) {
        i = pix_axis_ref(ny, cz, g.cam.ci);
        j = pix_axis_ref(nx, cz, g.cam.cj);
    }
    return ((unsigned)i < (unsigned)g.H && (unsigned)j < (unsigned)g.W) ? i * g.W + j : -1;
}

// project_voxel with the row and column: false when outside the image.
__device__ __forceinline__ bool project_voxel_ij(const GridView& g, float cx, float cy, float cz, int& i, int& j)
{
    const float r = __builtin_amdgcn_rcpf(cz);
    const float ny = cy * g.cam.fy, nx = cx * g.cam.fx;
    const bool ok = ((int)rcp_certified(r, cz) & (int)pix_axis_fast(ny, r, g.cam.ci, i) &
                     (int)pix_axis_fast(nx, r, g.cam.cj, j)) != 0;
    if (!ok) {
        i = pix_axis_ref(ny, cz, g.cam.ci);
        j = pix_axis_ref(nx, cz, g.cam.cj);
    }
    return (unsigned)i < (unsigned)g.H && (unsigned)j < (unsigned)g.W;
}

// The voxels (x, y, 0..7) of one lane, held in registers across ops.
template <bool kColor>
struct LaneVoxels {
    float s[kVPE], w[kVPE];
    float c[kColor ? kVPE : 1][3];
};

// One integrate (kSign = +1) or deintegrate (kSign = -1; kSign = 0: op.sign)
// op applied to a lane's eight voxels, in the reference's arithmetic order:
//   c = transformPoint(vm, E_inv), (row, col) = getCoordinate(c)
//   vpt = inputPcd(row, col) (from ptmap; x is NaN where z <= rmin, which
//   makes sdf NaN and fails the |sdf| < trunc test exactly as the
//   reference's `vpt.z > rmin` test does)
//   sdf = (vpt - vm) . Ez; update when |sdf| < trunc:
//     +1: s = (s w + sdf) / (w + 1), w = w + 1 (capped at max_weight, project
//         semantics: the reference never caps, GPUVoxelGrid.h:14,62)
//     -1: w != 1: s = (s w - sdf) / (w - 1), w = w - 1; else s = w = 0

// The update of voxel z of a lane (in the band of the op: pix >= 0 and
// |sdf| < trunc) from its register values.
template <bool kColor, int kSign>
__device__ __forceinline__ void update_voxel(const GridView& g, const OpDesc& op, int z, float sdf, int pix,
                                             LaneVoxels<kColor>& v)
{
    const bool use_rgb = kColor && op.rgb != nullptr;
    const float s = v.s[z], w = v.w[z];
    float px[3] = {0.f, 0.f, 0.f};
    if (use_rgb) {
        const uint8_t* q = op.rgb + (size_t)pix * 3;
        px[0] = (float)q[0];
        px[1] = (float)q[1];
        px[2] = (float)q[2];
    }
    if (kSign == 0) {
        // sign known only at run time (uniform per op): a - b is
        // computed as a + (-b), which IEEE arithmetic makes identical
        const bool pos = op.sign > 0;
        const float sg = pos ? 1.0f : -1.0f;
        const float d = w + sg;
        const bool zero = !pos && w == 1.0f;
        const bool cap = pos && g.wmax > 0.0f && w >= g.wmax;
        const float q = (s * w + sg * sdf) / d;
        v.s[z] = zero ? 0.0f : q;
        if (use_rgb)
            for (int ch = 0; ch < 3; ++ch)
                v.c[z][ch] = zero ? 0.0f : (v.c[z][ch] * w + sg * px[ch]) / d;
        v.w[z] = zero ? 0.0f : (cap ? g.wmax : d);
    } else if (kSign > 0) {
        const float d = w + 1.0f;
        v.s[z] = (s * w + sdf) / d;
        if (use_rgb)
            for (int ch = 0; ch < 3; ++ch)
                v.c[z][ch] = (v.c[z][ch] * w + px[ch]) / d;
        v.w[z] = (g.wmax > 0.0f && w >= g.wmax) ? g.wmax : d;
    } else {
        const float d = w - 1.0f;
        const bool one = w == 1.0f;
        v.s[z] = one ? 0.0f : (s * w - sdf) / d;
        if (use_rgb)
            for (int ch = 0; ch < 3; ++ch)
                v.c[z][ch] = one ? 0.0f : (v.c[z][ch] * w - px[ch]) / d;
        v.w[z] = one ? 0.0f : d;
    }
}

// The sdf and pixel of voxels h .. h + kN - 1 of a lane: the projections
// first, then the point-map gathers together, then the sdf.  Returns the
// band bits (pix >= 0 and |sdf| < trunc).
template <int kN>
__device__ __forceinline__ unsigned int op_band(const GridView& g, const OpDesc& op, float vmx, float vmy,
                                                const float (&vmz)[kVPE], int h, float (&sdf)[kVPE],
                                                int (&pix)[kVPE])
{
    const float* Ei = op.pose.Ei;
    // transformPoint with the z-independent part hoisted; the association
    // ((x*E0 + y*E1) + z*E2) + E3 is the reference's
    const float a0 = vmx * Ei[0] + vmy * Ei[1];
    const float a1 = vmx * Ei[4] + vmy * Ei[5];
    const float a2 = vmx * Ei[8] + vmy * Ei[9];
    const float ez0 = op.pose.Ez[0], ez1 = op.pose.Ez[1], ez2 = op.pose.Ez[2];
#pragma unroll
    for (int z = h; z < h + kN; ++z)
        pix[z] = project_voxel(g, a0 + vmz[z] * Ei[2] + Ei[3], a1 + vmz[z] * Ei[6] + Ei[7],
                               a2 + vmz[z] * Ei[10] + Ei[11]);
    float pt[kVPE][3];
#pragma unroll
    for (int z = h; z < h + kN; ++z) {
        const float* q = (const float*)(op.ptmap + (pix[z] >= 0 ? pix[z] : 0));
        pt[z][0] = q[0];
        pt[z][1] = q[1];
        pt[z][2] = q[2];
    }
    unsigned int band = 0;
#pragma unroll
    for (int z = h; z < h + kN; ++z) {
        sdf[z] = (pt[z][0] - vmx) * ez0 + (pt[z][1] - vmy) * ez1 + (pt[z][2] - vmz[z]) * ez2;
        band |= (pix[z] >= 0 && fabsf(sdf[z]) < g.trunc) ? 1u << z : 0u;
    }
    return band;
}

// op_band with the depth gathered instead of the point map: the world point
// of pixel (i, j) is rebuilt as getPoint and transformPoint make it
// (GPUVoxelGrid.h:229-232, 240-247): d = (float)depth / 1000 (RN32 through
// one double product, quot_rn), x = colfac[j] * d, y = rowfac[i] * d,
// z = d, then E; a pixel with d <= rmin is no measurement (the reference's
// `vpt.z > rmin`).  pix is i * W + j, or -1.
constexpr double kRmm = 0.001;  // RN64(1 / 1000)
template <int kN>
__device__ __forceinline__ unsigned int op_band_depth(const GridView& g, const OpDesc& op, float vmx, float vmy,
                                                      const float (&vmz)[kVPE], int h, float (&sdf)[kVPE],
                                                      int (&pix)[kVPE])
{
    const float* Ei = op.pose.Ei;
    const float* E = op.pose.E;
    const float a0 = vmx * Ei[0] + vmy * Ei[1];
    const float a1 = vmx * Ei[4] + vmy * Ei[5];
    const float a2 = vmx * Ei[8] + vmy * Ei[9];
    const float ez0 = op.pose.Ez[0], ez1 = op.pose.Ez[1], ez2 = op.pose.Ez[2];
    int ii[kVPE], jj[kVPE];
    bool in[kVPE];
#pragma unroll
    for (int z = h; z < h + kN; ++z)
        in[z] = project_voxel_ij(g, a0 + vmz[z] * Ei[2] + Ei[3], a1 + vmz[z] * Ei[6] + Ei[7],
                                 a2 + vmz[z] * Ei[10] + Ei[11], ii[z], jj[z]);
    uint16_t u[kVPE];
    float cf[kVPE], rf[kVPE];
#pragma unroll
    for (int z = h; z < h + kN; ++z) {
        const int i = in[z] ? ii[z] : 0, j = in[z] ? jj[z] : 0;
        u[z] = op.depth[(size_t)i * op.pitch + j];
        cf[z] = g.colfac[j];
        rf[z] = g.rowfac[i];
    }
    unsigned int band = 0;
#pragma unroll
    for (int z = h; z < h + kN; ++z) {
        const float d = quot_rn((float)u[z], kRmm);
        f3 p;
        p.x = cf[z] * d;
        p.y = rf[z] * d;
        p.z = d;
        const f3 pt = xform(p, E);
        sdf[z] = (pt.x - vmx) * ez0 + (pt.y - vmy) * ez1 + (pt.z - vmz[z]) * ez2;
        pix[z] = in[z] ? ii[z] * g.W + jj[z] : -1;
        band |= (in[z] && d > g.rmin && fabsf(sdf[z]) < g.trunc) ? 1u << z : 0u;
    }
    return band;
}

// A brick's rectangle of the point map staged in LDS: rows i0 .. i0 + h - 1,
// columns j0 .. j0 + w - 1, row-major; h = 0 when nothing is staged.
constexpr int kTilePx = 320;  // 5 KiB of LDS per wavefront
struct DepthTile {
    int i0, j0, h, w;
    const float4* px;
};

// Stages the point-map rectangle that brick (bx, by, bz) projects to.  Lanes
// 0..7 project the brick's 8 corners (an approximate reciprocal: the
// rectangle bounds the voxels' pixels with 2 pixels of margin, and a voxel
// whose exact pixel still falls outside it reads the point map in global
// memory, so no result depends on the rectangle).  One wavefront owns `lt`.
[[maybe_unused]] __device__ __forceinline__ DepthTile stage_depth_tile(const GridView& g, const OpDesc& op, int bx, int by, int bz,
                                                      int lane, float4* lt)
{
    const float* Ei = op.pose.Ei;
    const int c = lane & 7;
    const float px = g.vbl * (float)(bx + (c & 1)), py = g.vbl * (float)(by + ((c >> 1) & 1)),
                pz = g.vbl * (float)(bz + (c >> 2));
    const float cx = px * Ei[0] + py * Ei[1] + pz * Ei[2] + Ei[3];
    const float cy = px * Ei[4] + py * Ei[5] + pz * Ei[6] + Ei[7];
    const float cz = px * Ei[8] + py * Ei[9] + pz * Ei[10] + Ei[11];
    const float r = __builtin_amdgcn_rcpf(cz);
    // clamped to a range that converts to int exactly
    const float lim = 65536.0f;
    const float fi = fminf(fmaxf(cy * g.cam.fy * r + g.cam.ci, -lim), lim);
    const float fj = fminf(fmaxf(cx * g.cam.fx * r + g.cam.cj, -lim), lim);
    float imin = fi, imax = fi, jmin = fj, jmax = fj;
    bool front = cz > 1e-3f;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        imin = fminf(imin, __shfl_xor(imin, m));
        imax = fmaxf(imax, __shfl_xor(imax, m));
        jmin = fminf(jmin, __shfl_xor(jmin, m));
        jmax = fmaxf(jmax, __shfl_xor(jmax, m));
        front = front && __shfl_xor((int)front, m);
    }
    DepthTile t;
    t.px = lt;
    const float ui0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(imin)));
    const float ui1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(imax)));
    const float uj0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(jmin)));
    const float uj1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(jmax)));
    const bool ufront = __builtin_amdgcn_readfirstlane((int)front) != 0;
    t.i0 = max(0, (int)floorf(ui0) - 2);
    t.j0 = max(0, (int)floorf(uj0) - 2);
    const int i1 = min(g.H - 1, (int)floorf(ui1) + 2), j1 = min(g.W - 1, (int)floorf(uj1) + 2);
    t.h = i1 - t.i0 + 1;
    t.w = j1 - t.j0 + 1;
    if (!ufront || t.h <= 0 || t.w <= 0 || t.h * t.w > kTilePx) {
        t.h = 0;
        t.w = 1;
        return t;
    }
    const int area = t.h * t.w;
    for (int k = lane; k < area; k += 64) {
        const int row = k / t.w, col = k - row * t.w;
        lt[k] = op.ptmap[(size_t)(t.i0 + row) * g.W + (t.j0 + col)];
    }
    __builtin_amdgcn_wave_barrier();
    return t;
}

// op_band with the point map read from the brick's LDS rectangle where the
// voxel's exact pixel lies in it, from global memory elsewhere.
template <int kN>
__device__ __forceinline__ unsigned int op_band_tile(const GridView& g, const OpDesc& op, float vmx, float vmy,
                                                     const float (&vmz)[kVPE], int h, float (&sdf)[kVPE],
                                                     int (&pix)[kVPE], const DepthTile& dt)
{
    const float* Ei = op.pose.Ei;
    const float a0 = vmx * Ei[0] + vmy * Ei[1];
    const float a1 = vmx * Ei[4] + vmy * Ei[5];
    const float a2 = vmx * Ei[8] + vmy * Ei[9];
    const float ez0 = op.pose.Ez[0], ez1 = op.pose.Ez[1], ez2 = op.pose.Ez[2];
    int ii[kVPE], jj[kVPE];
    bool in[kVPE];
#pragma unroll
    for (int z = h; z < h + kN; ++z)
        in[z] = project_voxel_ij(g, a0 + vmz[z] * Ei[2] + Ei[3], a1 + vmz[z] * Ei[6] + Ei[7],
                                 a2 + vmz[z] * Ei[10] + Ei[11], ii[z], jj[z]);
    float4 q[kVPE];
#pragma unroll
    for (int z = h; z < h + kN; ++z) {
        const unsigned int di = (unsigned int)ii[z] - (unsigned int)dt.i0;
        const unsigned int dj = (unsigned int)jj[z] - (unsigned int)dt.j0;
        const bool inr = in[z] && di < (unsigned int)dt.h && dj < (unsigned int)dt.w;
        q[z] = dt.px[inr ? di * (unsigned int)dt.w + dj : 0u];
        if (in[z] && !inr)
            q[z] = op.ptmap[ii[z] * g.W + jj[z]];
    }
    unsigned int band = 0;
#pragma unroll
    for (int z = h; z < h + kN; ++z) {
        sdf[z] = (q[z].x - vmx) * ez0 + (q[z].y - vmy) * ez1 + (q[z].z - vmz[z]) * ez2;
        pix[z] = in[z] ? ii[z] * g.W + jj[z] : -1;
        band |= (in[z] && fabsf(sdf[z]) < g.trunc) ? 1u << z : 0u;
    }
    return band;
}

// The op applied to a lane's eight voxels held in registers, in two halves
// of kHalf voxels.  Returns a bit per voxel that was updated.
// kBand: 0 gathers the point map, 1 the depth (op_band_depth), 2 reads the
// brick's LDS rectangle dt (op_band_tile).
template <bool kColor, int kSign, int kBand = 0, int kHalf = 4>
__device__ __forceinline__ unsigned int apply_op(const GridView& g, const OpDesc& op, float vmx, float vmy,
                                                 const float (&vmz)[kVPE], LaneVoxels<kColor>& v,
                                                 const DepthTile& dt = DepthTile{0, 0, 0, 1, nullptr})
{
    unsigned int upd = 0;
#pragma unroll
    for (int h = 0; h < kVPE; h += kHalf) {
        float sdf[kVPE];
        int pix[kVPE];
        const unsigned int band = kBand == 1   ? op_band_depth<kHalf>(g, op, vmx, vmy, vmz, h, sdf, pix)
                                  : kBand == 2 ? op_band_tile<kHalf>(g, op, vmx, vmy, vmz, h, sdf, pix, dt)
                                               : op_band<kHalf>(g, op, vmx, vmy, vmz, h, sdf, pix);
#pragma unroll
        for (int z = h; z < h + kHalf; ++z)
            if ((band >> z) & 1u)
                update_voxel<kColor, kSign>(g, op, z, sdf[z], pix[z], v);
        upd |= band;
    }
    return upd;
}

template <typename T, bool kColor>
__device__ __forceinline__ void load_lane(const GridView& g, int id, int lane, bool fresh, LaneVoxels<kColor>& v)
{
    const T* sd = (const T*)g.payload + (size_t)id * (2 * kVPB);
    const float* cr = kColor ? g.rgb + (size_t)id * (3 * kVPB) : nullptr;
#pragma unroll
    for (int z = 0; z < kVPE; ++z) {
        v.s[z] = v.w[z] = 0.f;
        if (kColor)
            v.c[z][0] = v.c[z][1] = v.c[z][2] = 0.f;
        if (!fresh) {
            ld_pair(sd, lane + 64 * z, v.s[z], v.w[z]);
            if (kColor) {
                v.c[z][0] = cr[lane + 64 * z];
                v.c[z][1] = cr[lane + 64 * z + kVPB];
                v.c[z][2] = cr[lane + 64 * z + 2 * kVPB];
            }
        }
    }
}

// Writes back the voxels flagged in `dirty` (all of them for a fresh brick,
// whose payload memory was never initialised).  A 128-byte line of a z
// plane is written whole when any voxel in it changed (TSDF_STORE_LINES):
// the other lanes of the line store the values they loaded, which is exact
// because one wavefront owns the brick, and whole lines cost less HBM write
// traffic than partial ones.  All lanes of the wavefront must call this.
#ifndef TSDF_STORE_LINES
#define TSDF_STORE_LINES 1
#endif

__device__ __forceinline__ bool line_dirty(bool mine, int lane, int lanes_per_line)
{
    const unsigned long long b = __ballot(mine);
    const unsigned long long gm = lanes_per_line >= 64 ? ~0ull
                                                       : (((1ull << lanes_per_line) - 1ull) << (lane & ~(lanes_per_line - 1)));
    return (b & gm) != 0ull;
}

template <typename T, bool kColor>
__device__ __forceinline__ void store_lane(const GridView& g, int id, int lane, unsigned int dirty,
                                           const LaneVoxels<kColor>& v)
{
    T* sd = (T*)g.payload + (size_t)id * (2 * kVPB);
    float* cr = kColor ? g.rgb + (size_t)id * (3 * kVPB) : nullptr;
#pragma unroll
    for (int z = 0; z < kVPE; ++z) {
        const bool mine = (dirty >> z) & 1u;
        const bool wp = TSDF_STORE_LINES ? line_dirty(mine, lane, 128 / (2 * (int)sizeof(T))) : mine;
        const bool wc = kColor && (TSDF_STORE_LINES ? line_dirty(mine, lane, 32) : mine);
        const int k = lane + 64 * z;
        if (wp)
            st_pair(sd, k, v.s[z], v.w[z]);
        if (wc) {
            cr[k] = v.c[z][0];
            cr[k + kVPB] = v.c[z][1];
            cr[k + 2 * kVPB] = v.c[z][2];
        }
    }
}

// Voxel-centre coordinates of a lane: getVoxelMidPoint with the reference's
// double promotion of (k + 0.5) * vl (GPUVoxelGrid.cpp:501,521).
struct LaneMid {
    float x, y, z[kVPE];
};

__device__ __forceinline__ LaneMid lane_mid(const GridView& g, const Offsets& off, int bx, int by, int bz,
                                            int lane)
{
    LaneMid m;
    const float pzx = g.vbl * (float)bx, pzy = g.vbl * (float)by, pzz = g.vbl * (float)bz;
    m.x = (float)((double)pzx + off.v[lane & 7]);
    m.y = (float)((double)pzy + off.v[lane >> 3]);
#pragma unroll
    for (int z = 0; z < kVPE; ++z)
        m.z[z] = (float)((double)pzz + off.v[z]);
    return m;
}

// Which touched-list entries a wavefront handles.  Workgroups are dealt to
// the 8 XCDs round-robin; each XCD gets one contiguous eighth of the list so
// the bricks (and ptmap pixels) that one XCD's L2 serves are spatial
// neighbours.
struct WaveRange {
    unsigned int begin, end, stride;
};

__device__ __forceinline__ WaveRange wave_range(unsigned int count)
{
    const unsigned int xcd = blockIdx.x & 7u;
    const unsigned int per_xcd_wg = gridDim.x >> 3;  // gridDim.x is a multiple of 8
    const unsigned int local_wave = (blockIdx.x >> 3) * (kThreads / 64) + (threadIdx.x >> 6);
    const unsigned int seg = (count + 7u) >> 3;
    const unsigned int sb = min(count, xcd * seg), se = min(count, sb + seg);
    WaveRange r;
    r.begin = sb + local_wave;
    r.end = se;
    r.stride = per_xcd_wg * (kThreads / 64);
    return r;
}

// ---------------------------------------------------------------------------
// k_integrate: integrationCore[Color]Tag / deIntegrationCore[Color]Tag,
// GPUVoxelGrid.cpp:516-550.  One wavefront per touched brick, lane = (x, y),
// eight voxels along z per lane.
//
// A workgroup takes the touched list in chunks of kMetaChunk entries and
// first resolves slot -> brick id -> block index for the whole chunk at once
// (three dependent loads per chunk rather than per brick), then its four
// wavefronts update the chunk's bricks.  Within a brick, the eight payload
// loads and the eight depth-point gathers of a lane are issued together.
// ---------------------------------------------------------------------------
#ifndef TSDF_META_CHUNK
#define TSDF_META_CHUNK 8
#endif
constexpr int kMetaChunk = TSDF_META_CHUNK;  // touched entries resolved per workgroup round

// Adds every thread's `v` to statistics counter `id` with one atomic per
// workgroup, on the workgroup's lane (all threads of the workgroup must call
// it).  The wavefronts of a persistent update launch finish close together.
__device__ __forceinline__ void add_per_workgroup(const GridView& g, int id, unsigned int v)
{
    __shared__ unsigned int part[kThreads / 64];
    v = wave_sum(v);
    if (__lane_id() == 0)
        part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int k = 0; k < kThreads / 64; ++k)
            t += part[k];
        if (t)
            atomicAdd(stat_lane(g, id, workgroup_index()), (unsigned long long)t);
    }
}

// Resolves touched[t0 .. t1) (at most kMetaChunk entries) into LDS:
// (id | fresh << 31, x, y, z), id = -1 when the slot's allocation failed.
__device__ __forceinline__ void load_meta(const GridView& g, unsigned int t0, unsigned int t1, int4* meta)
{
    if (threadIdx.x < kMetaChunk) {
        const unsigned int t = t0 + threadIdx.x;
        int4 m = make_int4(-1, 0, 0, 0);
        if (t < t1) {
            const uint32_t id = (uint32_t)g.slots[g.touched[t]];
            if (id < g.n_bricks) {
                const int4 c = g.coords[id];
                m = make_int4((int)(id | ((uint32_t)(c.w & 1) << 31)), c.x, c.y, c.z);
                if (c.w & 1)
                    g.coords[id].w = c.w & ~1;  // the brick is written below
            }
        }
        meta[threadIdx.x] = m;
    }
}

template <typename T, bool kColor, int kSign>
__global__ __launch_bounds__(kThreads) void k_integrate(GridView g, OpDesc op, int par, Offsets off)
{
    __shared__ int4 meta[kMetaChunk];
#if TSDF_DEPTH_TILE
    __shared__ float4 tiles[kThreads / 64][kTilePx];
#endif
    const unsigned int count = g.ctrl->touched_count[par];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.ctrl->touched_count[par ^ 1] = 0u;  // ready for the next call
        g.ctrl->total_touched += count;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // each XCD walks one contiguous eighth of the touched list (spatial
    // neighbours share its L2); its workgroups stride over that eighth
    const unsigned int xcd = blockIdx.x & 7u, nwg = gridDim.x >> 3, wg = blockIdx.x >> 3;
    const unsigned int seg = (count + 7u) >> 3;
    const unsigned int sb = min(count, xcd * seg), se = min(count, sb + seg);
    unsigned int nupd = 0;
    for (unsigned int t0 = sb + wg * kMetaChunk; t0 < se; t0 += nwg * kMetaChunk) {
        __syncthreads();
        load_meta(g, t0, se, meta);
        __syncthreads();
        for (int e = wave; e < kMetaChunk; e += kThreads / 64) {
            const int4 mv = meta[e];
            const int4 m = make_int4(__builtin_amdgcn_readfirstlane(mv.x), __builtin_amdgcn_readfirstlane(mv.y),
                                     __builtin_amdgcn_readfirstlane(mv.z), __builtin_amdgcn_readfirstlane(mv.w));
            if (m.x == -1)
                continue;
            const bool fresh = m.x < 0;  // bit 31: allocated by this call
            const int id = m.x & 0x7FFFFFFF;
            LaneVoxels<kColor> v;
            load_lane<T, kColor>(g, id, lane, fresh, v);
            const LaneMid mid = lane_mid(g, off, m.y, m.z, m.w, lane);
#if TSDF_DEPTH_TILE
            const DepthTile dt = stage_depth_tile(g, op, m.y, m.z, m.w, lane, tiles[wave]);
            const unsigned int upd = apply_op<kColor, kSign, 2>(g, op, mid.x, mid.y, mid.z, v, dt);
#else
            const unsigned int upd =
                apply_op<kColor, kSign, TSDF_DEPTH_GATHER ? 1 : 0>(g, op, mid.x, mid.y, mid.z, v);
#endif
            nupd += __popc(upd);
            store_lane<T, kColor>(g, id, lane, fresh ? 0xFFu : upd, v);
        }
    }
    add_per_workgroup(g, kStUpdated, nupd);
}

// ---------------------------------------------------------------------------
// k_integrate_batch: every op of a batch applied to one brick while its
// voxels stay in registers; ops run in their original order, so the result
// equals the sequence of plain calls bit for bit.
// ---------------------------------------------------------------------------
template <typename T, bool kColor>
__global__ __launch_bounds__(kThreads) void k_integrate_batch(GridView g, const OpDesc* ops,
                                                              const uint32_t* epoch_base, int pass, Offsets off,
                                                              uint32_t* op_mask, int n_words)
{
    const int par = (int)((*epoch_base + (uint32_t)pass) & 1u);
    const unsigned int count = g.ctrl->touched_count[par];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.ctrl->touched_count[par ^ 1] = 0u;
        g.ctrl->total_touched += count;
    }
    const int lane = threadIdx.x & 63;
    unsigned int nupd = 0;
    const WaveRange wr = wave_range(count);
    // every value below that is the same for the whole wavefront is moved
    // to scalar registers (readfirstlane)
    for (unsigned int t = wr.begin; t < wr.end; t += wr.stride) {
        const uint32_t slot = __builtin_amdgcn_readfirstlane(g.touched[t]);
        const uint32_t id = __builtin_amdgcn_readfirstlane((uint32_t)g.slots[slot]);
        if (id >= g.n_bricks) {
            if (lane < n_words)
                op_mask[(size_t)slot * kMaskWords + lane] = 0u;
            continue;
        }
        const int4 c = g.coords[id];
        const int cw = __builtin_amdgcn_readfirstlane(c.w);
        const bool fresh = (cw & 1) != 0;
        if (fresh && lane == 0)
            g.coords[id].w = cw & ~1;
        LaneVoxels<kColor> v;
        load_lane<T, kColor>(g, (int)id, lane, fresh, v);
        const LaneMid mid = lane_mid(g, off, __builtin_amdgcn_readfirstlane(c.x), __builtin_amdgcn_readfirstlane(c.y),
                                     __builtin_amdgcn_readfirstlane(c.z), lane);
        unsigned int dirty = fresh ? 0xFFu : 0u;
        for (int k = 0; k < n_words; ++k) {
            uint32_t bits = __builtin_amdgcn_readfirstlane(op_mask[(size_t)slot * kMaskWords + k]);
            while (bits) {
                const int b = __ffs(bits) - 1;
                bits &= bits - 1u;
                const OpDesc& op = ops[k * 32 + b];
                const unsigned int u = apply_op<kColor, 0>(g, op, mid.x, mid.y, mid.z, v);
                dirty |= u;
                nupd += __popc(u);
            }
        }
        store_lane<T, kColor>(g, (int)id, lane, dirty, v);
        if (lane < n_words)  // every lane has read the mask words above
            op_mask[(size_t)slot * kMaskWords + lane] = 0u;
    }
    add_per_workgroup(g, kStUpdated, nupd);
}

// ---------------------------------------------------------------------------
// k_sbdf: the raycast's empty-space distance over superblocks (SbdfInfo),
// rebuilt only when Ctrl::sb_version moved since the last build.  One
// workgroup: the box of set superblocks is small (a 5 x 4 m room at 4 mm
// voxels is ~3.5 k superblocks); the passes are those of sbdf_build_host.
// ---------------------------------------------------------------------------
constexpr int kSbdfThreads = 1024;

__global__ __launch_bounds__(kSbdfThreads) void k_sbdf(GridView g, uint8_t* tmp)
{
    __shared__ int sh[8];  // lo[3], dim[3], valid, skip
    if (threadIdx.x == 0) {
        const unsigned int ver = g.ctrl->sb_version;
        const SbdfInfo inf = *g.sbdf_info;
        int valid = 1;
        int lo[3] = {0, 0, 0}, dim[3] = {0, 0, 0};
        if (g.ctrl->bbox_min[0] <= g.ctrl->bbox_max[0]) {
            size_t vol = 1;
            for (int a = 0; a < 3; ++a) {
                lo[a] = g.ctrl->bbox_min[a] >> 3;
                const int hi = g.ctrl->bbox_max[a] >> 3;
                dim[a] = hi - lo[a] + 1;
                valid = valid && lo[a] + kSbHalf >= 0 && hi + kSbHalf < kSbSide;
                vol *= (size_t)dim[a];
            }
            valid = valid && vol <= kSbdfCap;
        }
        for (int a = 0; a < 3; ++a) {
            sh[a] = lo[a];
            sh[3 + a] = dim[a];
        }
        sh[6] = valid;
        sh[7] = inf.valid && inf.version == ver;
        if (!sh[7] && (!valid || dim[0] == 0)) {
            SbdfInfo o;
            o.version = ver;
            o.valid = valid;
            for (int a = 0; a < 3; ++a) {
                o.lo[a] = lo[a];
                o.dim[a] = dim[a];
            }
            *g.sbdf_info = o;
            sh[7] = 1;  // nothing to compute
        }
    }
    __syncthreads();
    if (sh[7])
        return;
    const int lo[3] = {sh[0], sh[1], sh[2]};
    const int nx = sh[3], ny = sh[4], nz = sh[5];
    const int vol = nx * ny * nz;
    uint8_t* out = g.sbdf;
    // x: distance to the nearest set bit of the row, both directions
    for (int r = threadIdx.x; r < ny * nz; r += kSbdfThreads) {
        const int y = r % ny, z = r / ny;
        uint8_t* row = out + (size_t)nx * r;
        int d = kSbdfMax;
        for (int x = 0; x < nx; ++x) {
            d = sb_bit(g.sb_bits, x + lo[0], y + lo[1], z + lo[2]) ? 0 : (d < kSbdfMax ? d + 1 : kSbdfMax);
            row[x] = (uint8_t)d;
        }
        d = kSbdfMax;
        for (int x = nx - 1; x >= 0; --x) {
            const int v = row[x];
            d = v == 0 ? 0 : (d < kSbdfMax ? d + 1 : kSbdfMax);
            row[x] = (uint8_t)(v < d ? v : d);
        }
    }
    __syncthreads();
    // y, then z: min over the window of max(|offset|, value)
    for (int c = threadIdx.x; c < vol; c += kSbdfThreads) {
        const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
        int best = kSbdfMax;
        const int y0 = y - kSbdfMax > 0 ? y - kSbdfMax : 0, y1 = y + kSbdfMax < ny - 1 ? y + kSbdfMax : ny - 1;
        for (int yy = y0; yy <= y1; ++yy) {
            const int a = yy > y ? yy - y : y - yy;
            const int v = out[x + nx * (yy + ny * z)];
            const int m = a > v ? a : v;
            best = m < best ? m : best;
        }
        tmp[c] = (uint8_t)best;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < vol; c += kSbdfThreads) {
        const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
        int best = kSbdfMax;
        const int z0 = z - kSbdfMax > 0 ? z - kSbdfMax : 0, z1 = z + kSbdfMax < nz - 1 ? z + kSbdfMax : nz - 1;
        for (int zz = z0; zz <= z1; ++zz) {
            const int a = zz > z ? zz - z : z - zz;
            const int v = tmp[x + nx * (y + ny * zz)];
            const int m = a > v ? a : v;
            best = m < best ? m : best;
        }
        out[c] = (uint8_t)best;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        SbdfInfo o;
        o.version = g.ctrl->sb_version;
        o.valid = 1;
        for (int a = 0; a < 3; ++a) {
            o.lo[a] = lo[a];
            o.dim[a] = sh[3 + a];
        }
        *g.sbdf_info = o;
    }
}

// ---------------------------------------------------------------------------
// k_raycast: raycastingTag, GPUVoxelGrid.cpp:561-715 (march in tsdf_march.hpp).
// One pixel per lane, 8 x 8 pixels per wavefront; a workgroup of kRayWaves
// wavefronts takes a (8 kRayWx) x (8 kRayWy) tile.
// ---------------------------------------------------------------------------
#ifndef TSDF_RAY_WAVES
#define TSDF_RAY_WAVES 4
#endif
constexpr int kRayWaves = TSDF_RAY_WAVES;  // 1, 2 or 4
constexpr int kRayWx = kRayWaves >= 2 ? 2 : 1, kRayWy = kRayWaves >= 4 ? 2 : 1;
constexpr int kRayTileW = 8 * kRayWx, kRayTileH = 8 * kRayWy;

// The outputs of one ray at pixel o: the world point, its colour, its
// camera depth (0 unless a crossing) and its UB flag.
__device__ __forceinline__ void write_ray(const RayResult& r, size_t o, const Pose& pose, float* points, float* colors,
                                          float* depth, uint8_t* ub)
{
    points[o * 3 + 0] = r.p.x;
    points[o * 3 + 1] = r.p.y;
    points[o * 3 + 2] = r.p.z;
    if (colors) {
        colors[o * 3 + 0] = r.c.x;
        colors[o * 3 + 1] = r.c.y;
        colors[o * 3 + 2] = r.c.z;
    }
    if (depth)
        depth[o] = r.code == 1 ? (r.p.x * pose.Ei[8] + r.p.y * pose.Ei[9] + r.p.z * pose.Ei[10] + pose.Ei[11]) : 0.0f;
    if (ub)
        ub[o] = (uint8_t)(r.code < 0 ? -r.code : 0);
}

// k_raycast_tail: the rays k_raycast deferred (queue[0, ray_q[qpar])),
// kRayGroup lanes per ray (march_group).  Group k of the grid takes rays k,
// k + groups, ... (a shared counter would put every group's atomic on one
// address).  The first thread clears the other parity's count for the next
// raycast.
#ifndef TSDF_TAIL_WG_PER_CU
#define TSDF_TAIL_WG_PER_CU 8  // workgroups of the deferred-ray march per CU
#endif
#ifndef TSDF_RAY_GROUP
#define TSDF_RAY_GROUP 16
#endif
constexpr int kRayGroup = TSDF_RAY_GROUP;

template <typename T, bool kColor>
__global__ __launch_bounds__(kThreads) void k_raycast_tail(GridView g, Pose pose, RayParams rp, float* points,
                                                           float* colors, float* depth, uint8_t* ub,
                                                           const RayState* queue, int qpar)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        g.ctrl->ray_q[qpar ^ 1] = 0u;
    const Group<kRayGroup> gr;
    const unsigned int n = g.ctrl->ray_q[qpar];
    const unsigned int groups = gridDim.x * (kThreads / kRayGroup);
    unsigned int hit = 0, bad = 0, samples = 0, probes = 0;
    for (unsigned int e = blockIdx.x * (kThreads / kRayGroup) + threadIdx.x / kRayGroup; e < n; e += groups) {
        const RayState st = queue[e];
        const RayResult r = march_group<T, kColor, kRayGroup>(g, pose, rp, st, gr);
        if (gr.t == 0) {
            write_ray(r, st.pix, pose, points, colors, depth, ub);
            hit += r.code == 1;
            bad += r.code < 0;
            samples += r.samples;
            probes += r.probes;
        }
    }
    hit = wave_sum(hit);
    bad = wave_sum(bad);
    samples = wave_sum(samples);
    probes = wave_sum(probes);
    if (__lane_id() == 0) {
        const unsigned int who = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
        if (hit) atomicAdd(stat_lane(g, kStRayHits, who), (unsigned long long)hit);
        if (bad) atomicAdd(stat_lane(g, kStRayUb, who), (unsigned long long)bad);
        if (samples) atomicAdd(stat_lane(g, kStRaySamples, who), (unsigned long long)samples);
        if (probes) atomicAdd(stat_lane(g, kStRayProbes, who), (unsigned long long)probes);
    }
}

template <typename T, bool kColor, bool kStats>
__global__ __launch_bounds__(64 * kRayWaves)
#ifdef TSDF_RAY_WAVES_PER_EU
__attribute__((amdgpu_waves_per_eu(TSDF_RAY_WAVES_PER_EU, TSDF_RAY_WAVES_PER_EU)))
#endif
void k_raycast(GridView g, Pose pose, RayParams rp, float* points, float* colors, float* depth, uint8_t* ub,
               Footprint fp, TileBounds tb, RayState* queue, int qpar)
{
    // the wavefront's start on the 100 MHz constant clock (its duration
    // orders the next raycast's tiles)
    const unsigned long long t_begin = __builtin_amdgcn_s_memrealtime();
#if TSDF_RAY_CACHE
    __shared__ unsigned long long rcache[kRayCache];
    for (int e = threadIdx.x; e < kRayCache; e += blockDim.x)
        rcache[e] = (unsigned long long)(uint32_t)(e ^ 1) << 32;
    __syncthreads();
#endif
    const int tiles_x = (rp.W + kRayTileW - 1) / kRayTileW;
    const int n_tiles = tiles_x * ((rp.H + kRayTileH - 1) / kRayTileH);
    // Workgroups are dealt to the 8 XCDs round-robin; with xcd_map the
    // workgroups of one XCD take one contiguous band of tiles, so the bricks
    // one XCD's L2 holds belong to neighbouring rays.
    const int per = (n_tiles + 7) >> 3;
    const int b = (int)blockIdx.x;
    int tile = rp.xcd_map ? (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3) : b;
    // the dispatch order (the results do not depend on it): the last
    // workgroups dispatched start only when earlier ones finish, so a slow
    // tile dispatched late adds its whole time to the launch
    if (rp.order == 1 && b < n_tiles)
        tile = n_tiles - 1 - b;
    else if (rp.order == 2 && b < n_tiles)
        tile = (int)rp.tile_order[b];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = (tile / tiles_x) * kRayTileH + (wv / kRayWx) * 8 + (lane >> 3);
    const int j = (tile % tiles_x) * kRayTileW + (wv % kRayWx) * 8 + (lane & 7);
    // empty-space certificate for this raycast (identical in every lane):
    // probed blocks lie within `reach` blocks of the camera's block
    RayParams rq = rp;
#if TSDF_RAY_CACHE
    rq.lc = kStats ? nullptr : rcache;
#endif
    if (kStats) {
        rq.use_sb = 0;  // the footprint must see every probed key
        rq.use_tb = 0;
    } else if (rq.use_sb || rq.use_tb) {
        const int reach = (int)(rp.dmax / rp.vbl) + 3;
        const float* Ep = pose.E;
        int c[3] = {vb_index(Ep[3], rp.vbl), vb_index(Ep[7], rp.vbl), vb_index(Ep[11], rp.vbl)};
        int plo[3], phi[3], bmin[3], bmax[3];
        for (int a = 0; a < 3; ++a) {
            plo[a] = c[a] - reach;
            phi[a] = c[a] + reach;
            bmin[a] = g.ctrl->bbox_min[a];
            bmax[a] = g.ctrl->bbox_max[a];
        }
        const int cert = alias_free(bmin, bmax, plo, phi) ? 1 : 0;
        rq.use_sb = rq.use_sb && cert;
        rq.use_tb = rq.use_tb && cert;
        // and a distance field of the current superblock bits
        const SbdfInfo inf = rq.use_sb ? *g.sbdf_info : SbdfInfo{};
        if (rq.use_sb && inf.valid && inf.version == g.ctrl->sb_version) {
            rq.df = g.sbdf;
            for (int a = 0; a < 3; ++a) {
                rq.df_lo[a] = inf.lo[a];
                rq.df_dim[a] = inf.dim[a];
            }
        } else {
            rq.use_sb = 0;
        }
    }
    unsigned int hit = 0, bad = 0, samples = 0, probes = 0;
    // the depth bound of this wavefront's 8 x 8 pixels (one tile of the map)
    float zt = -1.0f;
    if (rq.use_tb && tile < n_tiles)
        zt = tb_depth(tb.z[(size_t)(i / kTbTile) * tb.tiles_x + (j / kTbTile)], tb.epoch);
    RayState st;
    bool defer = false;
    if (tile < n_tiles && i < rp.H && j < rp.W) {
        const RayResult r = march<T, kColor, kStats>(g, pose, rq, i, j, fp, zt, kStats ? 0 : rq.budget, &st);
        defer = r.code == kRayDeferred;
        if (!defer) {
            write_ray(r, (size_t)i * rp.W + j, pose, points, colors, depth, ub);
            hit = r.code == 1;
            bad = r.code < 0;
            samples = r.samples;
            probes = r.probes;
        }
    }
    // deferred rays: one append per wavefront; k_raycast_tail finishes them
    if (!kStats) {
        const unsigned long long dm = __ballot(defer);
        if (dm) {
            unsigned int base = 0;
            const int leader = __ffsll((long long)dm) - 1;
            if (lane == leader)
                base = atomicAdd(&g.ctrl->ray_q[qpar], (unsigned int)__popcll(dm));
            base = __shfl(base, leader);
            if (defer)
                queue[base + __popcll(dm & ((1ull << lane) - 1ull))] = st;
        }
    }
    hit = wave_sum(hit);
    bad = wave_sum(bad);
    samples = wave_sum(samples);
    probes = wave_sum(probes);
    // this wavefront's time, for the next raycast's schedule (it decides
    // only the order of the tiles, never a result)
    if (!kStats && rp.tile_cost && tile < n_tiles && lane == 0)
        rp.tile_cost[(size_t)tile * kRayWaves + wv] = (uint32_t)(__builtin_amdgcn_s_memrealtime() - t_begin);
    // one atomic per wavefront and counter, on the wavefront's lane: the
    // wavefronts end at very different times, so none waits for the others
    if (lane == 0) {
        const unsigned int who = blockIdx.x * kRayWaves + (unsigned int)wv;
        if (hit) atomicAdd(stat_lane(g, kStRayHits, who), (unsigned long long)hit);
        if (bad) atomicAdd(stat_lane(g, kStRayUb, who), (unsigned long long)bad);
        if (samples) atomicAdd(stat_lane(g, kStRaySamples, who), (unsigned long long)samples);
        if (probes) atomicAdd(stat_lane(g, kStRayProbes, who), (unsigned long long)probes);
    }
}

// k_ray_bounds: the tile depth bounds of one raycast (TileBounds in
// tsdf_march.hpp).  A workgroup takes a contiguous run of brick ids (bricks
// allocated together are spatial neighbours, so a run covers few tiles),
// lowers the bounds of the tiles each grown brick may cover in a copy of the
// tile map in LDS, then merges its copy into the global map with one atomic
// per tile it touched.  Larger images than kTbLdsTiles tiles lower the global
// map directly.
constexpr int kTbLdsTiles = 16384;  // 64 KiB of LDS: 1280 x 720 has 14 400 tiles
#ifndef TSDF_TB_PER
#define TSDF_TB_PER 128  // least bricks per workgroup run
#endif
#ifndef TSDF_TB_GLOBAL
#define TSDF_TB_GLOBAL 0
#endif
constexpr uint32_t kTbNone = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t f32_bits(float f)
{
    return __float_as_uint(f);
}

template <bool kLds>
__global__ __launch_bounds__(kThreads) void k_ray_bounds(GridView g, Pose P, TileBounds tb)
{
    extern __shared__ uint32_t lz[];  // [tiles] positive depths compare as their bits
    // the allocated count is known only here: runs of at least TSDF_TB_PER
    // bricks, the workgroups past the last run leave at once
    const unsigned int n = min(g.ctrl->alloc_top, g.n_bricks);
    const unsigned int per = max((n + gridDim.x - 1) / gridDim.x, (unsigned int)TSDF_TB_PER);
    const unsigned int b0 = blockIdx.x * per, b1 = min(n, b0 + per);
    if (b0 >= n)
        return;
    const int nt = tb.tiles_x * tb.tiles_y;
    if (kLds) {
        for (int t = threadIdx.x; t < nt; t += kThreads)
            lz[t] = kTbNone;
        __syncthreads();
    }
    for (unsigned int id = b0 + threadIdx.x; id < b1; id += kThreads) {
        const int4 c = g.coords[id];
        float zmin;
        int r0, r1, c0, c1;
        if (!brick_bounds(c.x, c.y, c.z, g.vbl, P, g.cam, g.H, g.W, tb.grow, tb.z_lo, zmin, r0, r1, c0, c1))
            continue;
        const uint32_t zb = f32_bits(zmin);
        for (int ty = r0 / kTbTile; ty <= r1 / kTbTile; ++ty)
            for (int tx = c0 / kTbTile; tx <= c1 / kTbTile; ++tx) {
                if (kLds) {
                    uint32_t* q = &lz[ty * tb.tiles_x + tx];
                    if (*q > zb)  // most bricks lie behind a nearer one of the tile
                        atomicMin(q, zb);
                }
                else
                    atomicMin(tb.z + (size_t)ty * tb.tiles_x + tx, tb_pack(tb.epoch, zmin));
            }
    }
    if (kLds) {
        __syncthreads();
        const unsigned long long hi = (unsigned long long)(~tb.epoch) << 32;
        for (int t = threadIdx.x; t < nt; t += kThreads) {
            const uint32_t v = lz[t];
            if (v != kTbNone)
                atomicMin(tb.z + t, hi | v);
        }
    }
}

// The slowest-first tile order for the next raycast: tiles sorted by the
// time of their slowest wavefront in this one, largest first (a counting
// sort over kOrderBuckets levels; the order within a level is arbitrary).
// One workgroup.
constexpr int kOrderThreads = 1024;
constexpr int kOrderBuckets = 256;
#ifndef TSDF_ORDER_EVERY
#define TSDF_ORDER_EVERY 4
#endif
constexpr int kOrderEvery = TSDF_ORDER_EVERY;  // raycasts per rebuild of the order
__global__ __launch_bounds__(kOrderThreads) void k_ray_order(const uint32_t* cost, int n_tiles, uint32_t* order)
{
    __shared__ unsigned int hist[kOrderBuckets];
    __shared__ unsigned int cmax;
    for (int k = threadIdx.x; k < kOrderBuckets; k += kOrderThreads)
        hist[k] = 0;
    if (threadIdx.x == 0)
        cmax = 0;
    __syncthreads();
    unsigned int m = 0;
    for (int t = threadIdx.x; t < n_tiles; t += kOrderThreads) {
        unsigned int c = 0;
        for (int w = 0; w < kRayWaves; ++w)
            c = max(c, cost[(size_t)t * kRayWaves + w]);
        m = max(m, c);
    }
    atomicMax(&cmax, m);
    __syncthreads();
    const unsigned long long top = (unsigned long long)cmax + 1ull;
    // level kOrderBuckets - 1 holds the slowest tiles
    for (int t = threadIdx.x; t < n_tiles; t += kOrderThreads) {
        unsigned int c = 0;
        for (int w = 0; w < kRayWaves; ++w)
            c = max(c, cost[(size_t)t * kRayWaves + w]);
        atomicAdd(&hist[(unsigned int)((unsigned long long)c * kOrderBuckets / top)], 1u);
    }
    __syncthreads();
    // exclusive offsets, slowest level first: a scan over the levels, one
    // per thread of the first kOrderBuckets threads (wavefront scans, then
    // the totals of the wavefronts before)
    __shared__ unsigned int wsum[kOrderBuckets / 64];
    unsigned int h = 0, incl = 0;
    const int k = kOrderBuckets - 1 - (int)threadIdx.x;
    if ((int)threadIdx.x < kOrderBuckets) {
        h = hist[k];
        incl = h;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int y = __shfl_up(incl, o);
            if ((int)__lane_id() >= o)
                incl += y;
        }
        if (__lane_id() == 63)
            wsum[threadIdx.x >> 6] = incl;
    }
    __syncthreads();
    if ((int)threadIdx.x < kOrderBuckets) {
        unsigned int pre = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); ++w)
            pre += wsum[w];
        hist[k] = pre + incl - h;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n_tiles; t += kOrderThreads) {
        unsigned int c = 0;
        for (int w = 0; w < kRayWaves; ++w)
            c = max(c, cost[(size_t)t * kRayWaves + w]);
        order[atomicAdd(&hist[(unsigned int)((unsigned long long)c * kOrderBuckets / top)], 1u)] = (uint32_t)t;
    }
}

__global__ __launch_bounds__(kThreads) void k_popcount(const uint32_t* bits, size_t n, unsigned long long* out)
{
    unsigned int c = 0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        c += __popc(bits[k]);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0 && c)
        atomicAdd(out, (unsigned long long)c);
}

// ---------------------------------------------------------------------------
// k_export: bricks -> caller buffers in fp32 (one workgroup per brick)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_export(GridView g, unsigned int n, int32_t* coords, float* sdf,
                                                     float* w, float* rgb)
{
    for (unsigned int b = blockIdx.x; b < n; b += gridDim.x) {
        const T* sd = (const T*)g.payload + (size_t)b * (2 * kVPB);
        if (threadIdx.x == 0 && coords) {
            const int4 c = g.coords[b];
            coords[(size_t)b * 3 + 0] = c.x;
            coords[(size_t)b * 3 + 1] = c.y;
            coords[(size_t)b * 3 + 2] = c.z;
        }
        const bool fresh = (g.coords[b].w & 1) != 0;  // allocated, never written
        for (int v = threadIdx.x; v < kVPB; v += kThreads) {
            float vs = 0.f, vw = 0.f;
            if (!fresh)
                ld_pair(sd, v, vs, vw);
            if (sdf) sdf[(size_t)b * kVPB + v] = vs;
            if (w) w[(size_t)b * kVPB + v] = vw;
            if (rgb && g.rgb)
                for (int ch = 0; ch < 3; ++ch)
                    rgb[((size_t)b * kVPB + v) * 3 + ch] = fresh ? 0.f : g.rgb[(size_t)b * 3 * kVPB + ch * kVPB + v];
        }
    }
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================

struct tsdf_grid {
    tsdf_params p;
    float vbl;
    bool intr_set;
    Camera cam;
    uint32_t epoch;
    Offsets off;
    int n_cu;
    // device buffers
    unsigned long long* slots;
    uint32_t* slot_epoch;
    int4* coords;
    void* payload;
    float* rgb;
    uint32_t* touched;
    uint32_t* sb_bits;
    uint64_t map_gen;    // bumped by every call that can set or clear superblock bits
    uint64_t sbdf_gen;   // map_gen at the last distance-field refresh
    uint8_t* sbdf;       // empty-space distance over superblocks, and its scratch
    uint8_t* sbdf_tmp;
    SbdfInfo* sbdf_info;
    float4* ptmap;
    float* colfac;  // [W] getPoint's column factors ((float)j - cj) / fx
    float* rowfac;  // [H] getPoint's row factors ((float)i - ci) / fy
    Ctrl* ctrl;
    Ctrl* ctrl_host;  // pinned
    unsigned long long* stat;       // statistics counters, kStatWords
    unsigned long long* stat_host;  // pinned
    unsigned long long* tb_z;       // [8x8-pixel tiles] depth bounds of the raycasts (TileBounds)
    RayState* ray_queue;            // [H * W] rays a raycast deferred to its tail kernel
    uint32_t ray_qepoch;            // raycasts that used the queue (their parity picks the counters)
    uint32_t tb_epoch;
    uint32_t* ray_cost;             // [raycast tiles * waves] work of the last raycast's wavefronts
    uint32_t* ray_order;            // [raycast tiles] slowest-first order built from it
    bool ray_order_valid;
    int ray_order_age;              // raycasts since the order was rebuilt
    // batch buffers (allocated on first batched call)
    uint32_t* op_mask;
    float4* ptmap_batch;
    BatchStage* stage_dev;
    BatchStage* stage_host;  // pinned
    // the batch as one captured graph, rebuilt when its shape changes
    hipStream_t cap_stream;
    hipGraphExec_t batch_exec;
    int batch_key[5];
    int batch_cap;     // point maps allocated in ptmap_batch
    hipEvent_t staging_done;  // recorded after the last copy out of ops_host
    bool staged;
    // raycast footprint buffers (allocated on first tsdf_raycast_footprint)
    uint32_t* fp_bits;
    unsigned long long* fp_keys;
    unsigned long long* fp_counts;
    float* fp_points;
    uint32_t fp_key_mask;
    // per-kernel event timing (tsdf_profile)
    bool prof_on;
    int prof_used, prof_cap;
    hipEvent_t* prof_ev;  // 2 per record
    int* prof_kind;
    double prof_ms[TSDF_N_KERNEL_KINDS];
    uint64_t prof_n[TSDF_N_KERNEL_KINDS];
    tsdf_tuning tune;
    // render views (differentiable render / tracking)
    ViewPose* view_dev;
    ViewPose* view_host;  // pinned
    hipEvent_t render_done;
    bool render_staged;
    // fixed-order reductions of the render adjoint and the tracking sums,
    // grown on demand: per-workgroup partial sums, the (voxel, value)
    // records of grad_sdf with their sorted copies, the sort's temporary
    float* red_part;
    size_t red_part_cap;  // floats
    uint32_t* rec_key[2];
    float* rec_val[2];
    size_t rec_cap;  // records
    void* sort_tmp;
    size_t sort_tmp_cap;  // bytes
};

namespace {

thread_local std::string t_err;

tsdf_status fail(tsdf_status s, const std::string& msg)
{
    t_err = msg;
    return s;
}

#define HIPC(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(TSDF_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

struct DeviceScope {
    int prev = -1;
    bool ok = true;
    explicit DeviceScope(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        if (prev != dev)
            ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceScope()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev)
            (void)hipSetDevice(prev);
    }
};

GridView view(const tsdf_grid* g)
{
    GridView v;
    v.slots = g->slots;
    v.slot_epoch = g->slot_epoch;
    v.coords = g->coords;
    v.payload = g->payload;
    v.rgb = g->rgb;
    v.touched = g->touched;
    v.sb_bits = g->sb_bits;
    v.sbdf = g->sbdf;
    v.sbdf_info = g->sbdf_info;
    v.ctrl = g->ctrl;
    v.stat = g->stat;
    v.mask = g->p.n_buckets - 1u;
    v.n_bricks = g->p.n_bricks;
    v.vl = g->p.voxel_length;
    v.vbl = g->vbl;
    v.rvbl = 1.0f / g->vbl;
    v.Rvbl = 1.0 / (double)g->vbl;
    v.trunc = g->p.truncation;
    v.rmin = g->p.min_depth;
    v.dmax = g->p.max_ray_length;
    v.wmax = g->p.max_weight;
    v.H = g->p.height;
    v.W = g->p.width;
    v.cam = g->cam;
    v.colfac = g->colfac;
    v.rowfac = g->rowfac;
    return v;
}

bool pose_ok(const float* E)
{
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(E[i]))
            return false;
    if (E[12] != 0.0f || E[13] != 0.0f || E[14] != 0.0f || E[15] != 1.0f)
        return false;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0;
            for (int r = 0; r < 3; ++r)
                s += (double)E[r * 4 + a] * (double)E[r * 4 + b];
            if (std::fabs(s - (a == b ? 1.0 : 0.0)) > 1e-4)
                return false;
        }
    for (int r = 0; r < 3; ++r)
        if (std::fabs(E[r * 4 + 3]) > 1e5f)
            return false;
    return true;
}

void make_pose(const float* E, Pose* P)
{
    float Ei[16];
    tsdf_pose_prepare(E, Ei, P->Ez);
    for (int k = 0; k < 12; ++k) {
        P->E[k] = E[k];
        P->Ei[k] = Ei[k];
    }
    P->pad = 0.f;
}

int persistent_grid(const tsdf_grid* g)
{
    // resident workgroups per CU (tunable), a multiple of the 8 XCDs
    int n = g->n_cu * g->tune.integrate_wg_per_cu;
    n = (n + 7) & ~7;
    return n < 64 ? 64 : n;
}

template <typename T>
void launch_integrate(const tsdf_grid* g, const OpDesc& op, int par, hipStream_t s, bool color)
{
    const int grid = persistent_grid(g);
    const GridView v = view(g);
    if (color && op.sign > 0)
        hipLaunchKernelGGL((k_integrate<T, true, 1>), dim3(grid), dim3(kThreads), 0, s, v, op, par, g->off);
    else if (color)
        hipLaunchKernelGGL((k_integrate<T, true, -1>), dim3(grid), dim3(kThreads), 0, s, v, op, par, g->off);
    else if (op.sign > 0)
        hipLaunchKernelGGL((k_integrate<T, false, 1>), dim3(grid), dim3(kThreads), 0, s, v, op, par, g->off);
    else
        hipLaunchKernelGGL((k_integrate<T, false, -1>), dim3(grid), dim3(kThreads), 0, s, v, op, par, g->off);
}

template <typename T>
void launch_integrate_batch(const tsdf_grid* g, const OpDesc* ops, int n_ops, const uint32_t* epoch_base, int pass,
                            hipStream_t s, bool color)
{
    const int grid = persistent_grid(g);
    const int words = (n_ops + 31) / 32;
    if (color)
        hipLaunchKernelGGL((k_integrate_batch<T, true>), dim3(grid), dim3(kThreads), 0, s, view(g), ops, epoch_base,
                           pass, g->off, g->op_mask, words);
    else
        hipLaunchKernelGGL((k_integrate_batch<T, false>), dim3(grid), dim3(kThreads), 0, s, view(g), ops,
                           epoch_base, pass, g->off, g->op_mask, words);
}

tsdf_status check_image_args(const tsdf_grid* g, const uint16_t* depth, int32_t pitch)
{
    if (!depth)
        return fail(TSDF_BAD_ARG, "depth is NULL");
    if (pitch < g->p.width)
        return fail(TSDF_BAD_DIMS, "pitch smaller than the grid width");
    return TSDF_OK;
}

// Device buffers of a batch of n_ops ops: the per-slot op masks and op
// descriptors (once), and one point map per op of a pass (grown to
// min(n_ops, kMaxOps) maps; a batch never allocates once that is reserved).
tsdf_status ensure_batch_buffers(tsdf_grid* g, int n_ops)
{
    if (!g->op_mask) {
        const size_t nb = g->p.n_buckets;
        HIPC(hipMalloc(&g->op_mask, nb * kMaskWords * sizeof(uint32_t)));
        HIPC(hipMemset(g->op_mask, 0, nb * kMaskWords * sizeof(uint32_t)));
        HIPC(hipMalloc(&g->stage_dev, sizeof(BatchStage)));
        HIPC(hipHostMalloc(&g->stage_host, sizeof(BatchStage), hipHostMallocDefault));
        HIPC(hipEventCreateWithFlags(&g->staging_done, hipEventDisableTiming));
    }
    const int want = n_ops < kMaxOps ? n_ops : kMaxOps;
    if (g->batch_cap < want) {
        HIPC(hipDeviceSynchronize());  // a previous pass may still read the old maps
        (void)hipFree(g->ptmap_batch);
        g->ptmap_batch = nullptr;
        g->batch_cap = 0;
        HIPC(hipMalloc(&g->ptmap_batch, (size_t)want * g->p.height * g->p.width * sizeof(float4)));
        g->batch_cap = want;
    }
    HIPC(hipDeviceSynchronize());
    return TSDF_OK;
}

// Per-kernel timing: an event pair around each launch on its own stream.
tsdf_status prof_flush(tsdf_grid* g)
{
    if (g->prof_used == 0)
        return TSDF_OK;
    HIPC(hipEventSynchronize(g->prof_ev[2 * g->prof_used - 1]));
    for (int k = 0; k < g->prof_used; ++k) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, g->prof_ev[2 * k], g->prof_ev[2 * k + 1]));
        g->prof_ms[g->prof_kind[k]] += ms;
        g->prof_n[g->prof_kind[k]] += 1;
    }
    g->prof_used = 0;
    return TSDF_OK;
}

struct ProfScope {
    tsdf_grid* g;
    hipStream_t s;
    int slot = -1;
    ProfScope(tsdf_grid* g_, hipStream_t s_, int kind, bool enable = true) : g(g_), s(s_)
    {
        if (!g->prof_on || !enable)
            return;
        if (g->prof_used == g->prof_cap) {
            const int nc = g->prof_cap ? 2 * g->prof_cap : 1024;
            hipEvent_t* ev = (hipEvent_t*)realloc(g->prof_ev, 2 * (size_t)nc * sizeof(hipEvent_t));
            int* kd = (int*)realloc(g->prof_kind, (size_t)nc * sizeof(int));
            if (!ev || !kd) {
                if (ev) g->prof_ev = ev;
                if (kd) g->prof_kind = kd;
                return;
            }
            g->prof_ev = ev;
            g->prof_kind = kd;
            for (int k = 2 * g->prof_cap; k < 2 * nc; ++k)
                if (hipEventCreate(&g->prof_ev[k]) != hipSuccess)
                    return;
            g->prof_cap = nc;
        }
        slot = g->prof_used++;
        g->prof_kind[slot] = kind;
        (void)hipEventRecord(g->prof_ev[2 * slot], s);
    }
    ~ProfScope()
    {
        if (slot >= 0)
            (void)hipEventRecord(g->prof_ev[2 * slot + 1], s);
    }
};

// The tile-bounds view of one raycast.  `grow` bounds the distance between a
// march position and the exact point of its ray parameter (the float
// position rv drifts by at most half a float spacing per update and
// component, the ray length l likewise, the step vectors carry their own
// rounding), with the number of position updates of one march bounded by
// 4 (d_max + vbl) / min(vl, d_trunc) + 64: every outer round of the march
// advances l by at least min(vl, d_trunc) net, and every position update but
// one back-off per round advances it.
TileBounds tile_bounds(tsdf_grid* g, const Pose& P)
{
    TileBounds tb;
    tb.z = g->tb_z;
    tb.tiles_x = (g->p.width + kTbTile - 1) / kTbTile;
    tb.tiles_y = (g->p.height + kTbTile - 1) / kTbTile;
    tb.epoch = g->tb_epoch;
    const double vl = g->p.voxel_length, dt = g->p.truncation, vbl = g->vbl, dmax = g->p.max_ray_length;
    const double n_upd = 4.0 * (dmax + vbl) / (vl < dt ? vl : dt) + 64.0;
    double tmax = 0.0;
    for (int a = 0; a < 3; ++a)
        tmax = std::fmax(tmax, std::fabs((double)P.E[4 * a + 3]));
    auto spacing = [](double x) { return std::ldexp(1.0, std::ilogb(x) - 23); };  // float spacing at |x|
    const double R = tmax + dmax + vbl + 1.0;
    const double drift = std::sqrt(3.0) * n_upd * spacing(R) / 2.0 + n_upd * spacing(dmax + vbl) / 2.0 +
                         std::sqrt(3.0) * n_upd * spacing(vbl) / 2.0 + std::sqrt(3.0) * dmax * spacing(R) +
                         8.0 * spacing(R) + 1e-4;
    tb.grow = (float)drift;
    // the least depth per unit ray length over the image (its corners)
    const double ex = std::fmax(std::fabs((0.0 - g->cam.cj) / g->cam.fx), std::fabs((g->p.width - 1.0 - g->cam.cj) / g->cam.fx));
    const double ey = std::fmax(std::fabs((0.0 - g->cam.ci) / g->cam.fy), std::fabs((g->p.height - 1.0 - g->cam.ci) / g->cam.fy));
    const double nz_min = 1.0 / std::sqrt(1.0 + ex * ex + ey * ey);
    tb.z_lo = (float)(0.1 * nz_min * (1.0 - 1e-4) - drift);
    return tb;
}

template <bool kStats>
void launch_raycast(tsdf_grid* g, const Pose& P, float* points, float* colors, float* depth, uint8_t* ub,
                    const Footprint& fp, hipStream_t s)
{
    RayParams rp;
    rp.vl = g->p.voxel_length;
    rp.vbl = g->vbl;
    rp.dt = g->p.truncation;
    rp.dmax = g->p.max_ray_length;
    rp.Rvl = 1.0 / (double)rp.vl;
    rp.Rvbl = 1.0 / (double)rp.vbl;
    rp.back = rp.vbl - rp.dt;
    rp.H = g->p.height;
    rp.W = g->p.width;
    rp.xcd_map = g->tune.raycast_xcd_map;
    rp.order = kStats ? 0 : g->tune.raycast_order;
    rp.tile_order = g->ray_order;
    rp.tile_cost = (!kStats && rp.order == 2) ? g->ray_cost : nullptr;
    if (rp.order == 2 && !g->ray_order_valid)
        rp.order = 0;  // no previous raycast to order by yet
    rp.use_sb = g->tune.raycast_skip == 1;
    rp.use_tb = g->tune.raycast_skip == 2 && !kStats;
    // the group march replays no superblock distance field
    rp.budget = (kStats || rp.use_sb) ? 0 : g->tune.raycast_budget;
    rp.lc = nullptr;  // the kernel points it at its LDS cache (TSDF_RAY_CACHE)
    const int qpar = (int)(g->ray_qepoch & 1u);
    if (rp.budget > 0)
        ++g->ray_qepoch;
    rp.df = nullptr;
    for (int a = 0; a < 3; ++a)
        rp.df_lo[a] = rp.df_dim[a] = 0;
    if (rp.use_sb && !kStats && g->map_gen != g->sbdf_gen) {
        // a call since the last check may have set superblock bits: refresh
        // the distance field if they moved (the kernel compares versions)
        hipLaunchKernelGGL(k_sbdf, dim3(1), dim3(kSbdfThreads), 0, s, view(g), g->sbdf_tmp);
        g->sbdf_gen = g->map_gen;
    }
    TileBounds tb = tile_bounds(g, P);
    if (!(tb.z_lo > 1e-3f))
        rp.use_tb = 0;  // the drift bound leaves no depth to clip at
    if (rp.use_tb) {
        tb.epoch = ++g->tb_epoch;
        const int nt = tb.tiles_x * tb.tiles_y;
        if (nt <= kTbLdsTiles && !TSDF_TB_GLOBAL)
            hipLaunchKernelGGL((k_ray_bounds<true>), dim3(g->n_cu * 4), dim3(kThreads), (size_t)nt * sizeof(uint32_t),
                               s, view(g), P, tb);
        else
            hipLaunchKernelGGL((k_ray_bounds<false>), dim3(g->n_cu * 4), dim3(kThreads), 0, s, view(g), P, tb);
    }
    const int nt = ((rp.W + kRayTileW - 1) / kRayTileW) * ((rp.H + kRayTileH - 1) / kRayTileH);
    const dim3 grid(((nt + 7) / 8) * 8);
    const dim3 block(64 * kRayWaves);
    const bool color = g->p.with_color != 0;
    if (g->p.storage == TSDF_STORAGE_F16) {
        if (color)
            hipLaunchKernelGGL((k_raycast<__half, true, kStats>), grid, block, 0, s, view(g), P, rp, points,
                               colors, depth, ub, fp, tb, g->ray_queue, qpar);
        else
            hipLaunchKernelGGL((k_raycast<__half, false, kStats>), grid, block, 0, s, view(g), P, rp,
                               points, colors, depth, ub, fp, tb, g->ray_queue, qpar);
    } else {
        if (color)
            hipLaunchKernelGGL((k_raycast<float, true, kStats>), grid, block, 0, s, view(g), P, rp, points,
                               colors, depth, ub, fp, tb, g->ray_queue, qpar);
        else
            hipLaunchKernelGGL((k_raycast<float, false, kStats>), grid, block, 0, s, view(g), P, rp, points,
                               colors, depth, ub, fp, tb, g->ray_queue, qpar);
    }
    if (rp.budget > 0) {
        // the deferred rays, kRayGroup lanes each (the queue length is read
        // on the device; groups past its end leave at once)
        const dim3 tgrid(g->n_cu * TSDF_TAIL_WG_PER_CU);
        if (g->p.storage == TSDF_STORAGE_F16) {
            if (color)
                hipLaunchKernelGGL((k_raycast_tail<__half, true>), tgrid, dim3(kThreads), 0, s, view(g), P, rp,
                                   points, colors, depth, ub, g->ray_queue, qpar);
            else
                hipLaunchKernelGGL((k_raycast_tail<__half, false>), tgrid, dim3(kThreads), 0, s, view(g), P, rp,
                                   points, colors, depth, ub, g->ray_queue, qpar);
        } else {
            if (color)
                hipLaunchKernelGGL((k_raycast_tail<float, true>), tgrid, dim3(kThreads), 0, s, view(g), P, rp,
                                   points, colors, depth, ub, g->ray_queue, qpar);
            else
                hipLaunchKernelGGL((k_raycast_tail<float, false>), tgrid, dim3(kThreads), 0, s, view(g), P, rp,
                                   points, colors, depth, ub, g->ray_queue, qpar);
        }
    }
    // the slowest-first order, rebuilt every kOrderEvery raycasts from the
    // last one's times (consecutive poses see nearly the same rays)
    if (rp.tile_cost && (!g->ray_order_valid || g->ray_order_age + 1 >= kOrderEvery)) {
        hipLaunchKernelGGL(k_ray_order, dim3(1), dim3(kOrderThreads), 0, s, g->ray_cost, nt, g->ray_order);
        g->ray_order_valid = true;
        g->ray_order_age = 0;
    } else if (rp.tile_cost) {
        ++g->ray_order_age;
    }
}

}  // namespace


extern "C" {


void tsdf_default_params(tsdf_params* p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->voxel_length = 0.005f;
    p->truncation = 0.02f;
    p->height = 480;
    p->width = 640;
    p->n_buckets = 1u << 20;
    p->n_bricks = 1u << 19;
    p->storage = TSDF_STORAGE_F32;
    p->with_color = 0;
    p->max_weight = 0.0f;
    p->min_depth = 0.2f;
    p->max_ray_length = 10.0f;
    p->device = 0;
}

const char* tsdf_status_string(tsdf_status s)
{
    switch (s) {
    case TSDF_OK: return "ok";
    case TSDF_BAD_ARG: return "bad argument";
    case TSDF_BAD_DIMS: return "image dimensions do not match the grid";
    case TSDF_MAP_FULL: return "brick pool or hash table full";
    case TSDF_HIP_ERROR: return "HIP runtime error";
    case TSDF_NO_INTRINSICS: return "intrinsics not set";
    case TSDF_BAD_POSE: return "pose is not a finite rigid transform";
    }
    return "unknown status";
}

const char* tsdf_last_error(void) { return t_err.c_str(); }

tsdf_status tsdf_pose_prepare(const float E[16], float Einv[16], float Ez[3])
{
    if (!E || !Einv || !Ez)
        return fail(TSDF_BAD_ARG, "NULL pose pointer");
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            Einv[r * 4 + c] = E[c * 4 + r];
        const double t = (double)E[0 * 4 + r] * (double)E[3] + (double)E[1 * 4 + r] * (double)E[7] +
                         (double)E[2 * 4 + r] * (double)E[11];
        Einv[r * 4 + 3] = (float)(-t);
    }
    Einv[12] = 0.f;
    Einv[13] = 0.f;
    Einv[14] = 0.f;
    Einv[15] = 1.f;
    const float x = E[2], y = E[6], z = E[10];
    const float len = std::sqrt(x * x + y * y + z * z);
    Ez[0] = x / len;
    Ez[1] = y / len;
    Ez[2] = z / len;
    return TSDF_OK;
}

tsdf_status tsdf_create(const tsdf_params* p, tsdf_grid** out)
{
    if (!p || !out)
        return fail(TSDF_BAD_ARG, "NULL params or output handle");
    *out = nullptr;
    const float vl = p->voxel_length, tr = p->truncation;
    if (!(vl >= 1e-4f && vl <= 1.0f))
        return fail(TSDF_BAD_ARG, "voxel_length must lie in [1e-4, 1] m");
    if (!(tr >= 1e-4f && tr <= 4.0f * kVPE * vl))
        return fail(TSDF_BAD_ARG, "truncation must lie in [1e-4 m, 4 brick lengths]");
    if (p->height < 1 || p->width < 1 || p->height > 16384 || p->width > 16384)
        return fail(TSDF_BAD_DIMS, "image size out of range");
    if (p->n_buckets < 1024u || (p->n_buckets & (p->n_buckets - 1u)) || p->n_buckets > (1u << 30))
        return fail(TSDF_BAD_ARG, "n_buckets must be a power of two in [2^10, 2^30]");
    if (p->n_bricks < 1u || p->n_bricks > p->n_buckets || p->n_bricks >= 0x80000000u)
        return fail(TSDF_BAD_ARG, "n_bricks must lie in [1, n_buckets]");
    if (p->storage != TSDF_STORAGE_F32 && p->storage != TSDF_STORAGE_F16)
        return fail(TSDF_BAD_ARG, "unknown storage");
    if (!(p->max_ray_length > 0.0f && p->max_ray_length <= 1000.0f) ||
        p->max_ray_length / (vl < tr ? vl : tr) > 1e6f)
        return fail(TSDF_BAD_ARG, "max_ray_length out of range");
    if (!(p->min_depth >= 0.0f) || !(p->max_weight >= 0.0f))
        return fail(TSDF_BAD_ARG, "min_depth and max_weight must be >= 0");
    if (p->storage == TSDF_STORAGE_F16 && p->max_weight > kF16MaxWeight)
        return fail(TSDF_BAD_ARG, "fp16 weights are exact only up to 2048");

    DeviceScope ds(p->device);
    if (!ds.ok)
        return fail(TSDF_HIP_ERROR, "hipSetDevice failed");
    tsdf_grid* g = new tsdf_grid();
    memset(g, 0, sizeof(*g));
    g->p = *p;
    // fp16 holds every integer weight up to 2048 and no larger one (2049
    // rounds to 2048), so an uncapped fp16 grid is capped there
    if (p->storage == TSDF_STORAGE_F16 && p->max_weight == 0.0f)
        g->p.max_weight = kF16MaxWeight;
    g->vbl = (float)kVPE * vl;  // GPUVoxelGrid.cpp:15
    for (int k = 0; k < kVPE; ++k)
        g->off.v[k] = ((double)k + 0.5) * (double)vl;
    hipDeviceProp_t prop;
    int ncu = 256;
    if (hipGetDeviceProperties(&prop, p->device) == hipSuccess)
        ncu = prop.multiProcessorCount;
    g->n_cu = ncu;
    tsdf_default_tuning(&g->tune);

    const size_t nb = p->n_buckets, nbr = p->n_bricks;
    const size_t tsz = p->storage == TSDF_STORAGE_F16 ? 2 : 4;
    tsdf_status st = TSDF_OK;
    auto alloc = [&](void** ptr, size_t bytes) {
        if (st != TSDF_OK)
            return;
        hipError_t e = hipMalloc(ptr, bytes);
        if (e != hipSuccess)
            st = fail(TSDF_HIP_ERROR, std::string("hipMalloc: ") + hipGetErrorString(e));
    };
    alloc((void**)&g->slots, nb * sizeof(unsigned long long));
    alloc((void**)&g->slot_epoch, nb * sizeof(uint32_t));
    alloc((void**)&g->coords, nbr * sizeof(int4));
    alloc(&g->payload, nbr * 2 * kVPB * tsz);
    if (p->with_color)
        alloc((void**)&g->rgb, nbr * 3 * kVPB * sizeof(float));
    alloc((void**)&g->touched, nb * sizeof(uint32_t));
    alloc((void**)&g->sb_bits, kSbWords * sizeof(uint32_t));
    alloc((void**)&g->sbdf, kSbdfCap);
    alloc((void**)&g->sbdf_tmp, kSbdfCap);
    alloc((void**)&g->sbdf_info, sizeof(SbdfInfo));
    alloc((void**)&g->ptmap, (size_t)p->height * p->width * sizeof(float4));
    alloc((void**)&g->colfac, (size_t)p->width * sizeof(float));
    alloc((void**)&g->rowfac, (size_t)p->height * sizeof(float));
    alloc((void**)&g->ctrl, sizeof(Ctrl));
    alloc((void**)&g->stat, kStatWords * sizeof(unsigned long long));
    {
        const size_t tiles = (size_t)((p->width + kRayTileW - 1) / kRayTileW) * ((p->height + kRayTileH - 1) / kRayTileH);
        alloc((void**)&g->ray_cost, tiles * kRayWaves * sizeof(uint32_t));
        alloc((void**)&g->ray_order, tiles * sizeof(uint32_t));
        const size_t t8 = (size_t)((p->width + kTbTile - 1) / kTbTile) * ((p->height + kTbTile - 1) / kTbTile);
        alloc((void**)&g->tb_z, t8 * sizeof(unsigned long long));
        alloc((void**)&g->ray_queue, (size_t)p->width * p->height * sizeof(RayState));
        if (st == TSDF_OK) {
            // no entry carries an epoch yet (epochs start at 1, ~1 != ~0)
            hipError_t e = hipMemset(g->tb_z, 0xFF, t8 * sizeof(unsigned long long));
            if (e != hipSuccess)
                st = fail(TSDF_HIP_ERROR, std::string("hipMemset: ") + hipGetErrorString(e));
        }
    }
    if (st == TSDF_OK) {
        // no distance field yet; counters start from zero
        hipError_t e = hipMemset(g->sbdf_info, 0, sizeof(SbdfInfo));
        if (e == hipSuccess)
            e = hipMemset(g->ctrl, 0, sizeof(Ctrl));
        if (e == hipSuccess)
            e = hipMemset(g->stat, 0, kStatWords * sizeof(unsigned long long));
        if (e != hipSuccess)
            st = fail(TSDF_HIP_ERROR, std::string("hipMemset: ") + hipGetErrorString(e));
    }
    alloc((void**)&g->view_dev, 256 * sizeof(ViewPose));
    if (st == TSDF_OK) {
        hipError_t e = hipHostMalloc((void**)&g->ctrl_host, sizeof(Ctrl), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&g->stat_host, kStatWords * sizeof(unsigned long long), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&g->view_host, 256 * sizeof(ViewPose), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&g->render_done, hipEventDisableTiming);
        if (e != hipSuccess)
            st = fail(TSDF_HIP_ERROR, std::string("hipHostMalloc / hipEventCreate: ") + hipGetErrorString(e));
    }
    if (st == TSDF_OK)
        st = tsdf_reset(g, nullptr);
    if (st == TSDF_OK && hipStreamSynchronize(nullptr) != hipSuccess)
        st = fail(TSDF_HIP_ERROR, "synchronize after reset failed");
    if (st != TSDF_OK) {
        std::string keep = t_err;
        tsdf_destroy(g);
        t_err = keep;
        return st;
    }
    *out = g;
    return TSDF_OK;
}

void tsdf_destroy(tsdf_grid* g)
{
    if (!g)
        return;
    DeviceScope ds(g->p.device);
    (void)hipDeviceSynchronize();
    (void)hipFree(g->slots);
    (void)hipFree(g->slot_epoch);
    (void)hipFree(g->coords);
    (void)hipFree(g->payload);
    (void)hipFree(g->rgb);
    (void)hipFree(g->touched);
    (void)hipFree(g->sb_bits);
    (void)hipFree(g->sbdf);
    (void)hipFree(g->sbdf_tmp);
    (void)hipFree(g->sbdf_info);
    (void)hipFree(g->ptmap);
    (void)hipFree(g->colfac);
    (void)hipFree(g->rowfac);
    (void)hipFree(g->ctrl);
    (void)hipFree(g->stat);
    (void)hipFree(g->ray_cost);
    (void)hipFree(g->ray_order);
    (void)hipFree(g->tb_z);
    (void)hipFree(g->ray_queue);
    (void)hipFree(g->op_mask);
    (void)hipFree(g->ptmap_batch);
    if (g->batch_exec)
        (void)hipGraphExecDestroy(g->batch_exec);
    if (g->cap_stream)
        (void)hipStreamDestroy(g->cap_stream);
    (void)hipFree(g->stage_dev);
    (void)hipFree(g->fp_bits);
    (void)hipFree(g->fp_keys);
    (void)hipFree(g->fp_counts);
    (void)hipFree(g->fp_points);
    for (int k = 0; k < 2 * g->prof_cap; ++k)
        (void)hipEventDestroy(g->prof_ev[k]);
    free(g->prof_ev);
    free(g->prof_kind);
    if (g->ctrl_host)
        (void)hipHostFree(g->ctrl_host);
    if (g->stat_host)
        (void)hipHostFree(g->stat_host);
    if (g->view_host)
        (void)hipHostFree(g->view_host);
    if (g->render_done)
        (void)hipEventDestroy(g->render_done);
    (void)hipFree(g->view_dev);
    (void)hipFree(g->red_part);
    for (int k = 0; k < 2; ++k) {
        (void)hipFree(g->rec_key[k]);
        (void)hipFree(g->rec_val[k]);
    }
    (void)hipFree(g->sort_tmp);
    if (g->stage_host)
        (void)hipHostFree(g->stage_host);
    if (g->staging_done)
        (void)hipEventDestroy(g->staging_done);
    delete g;
}

tsdf_status tsdf_reset(tsdf_grid* g, void* stream)
{
    if (!g)
        return fail(TSDF_BAD_ARG, "NULL grid");
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    const size_t words = g->op_mask ? (size_t)g->p.n_buckets * kMaskWords : 0;
    {
        ProfScope ps(g, s, TSDF_K_RESET);
        hipLaunchKernelGGL(k_reset, dim3(g->n_cu * 4), dim3(kThreads), 0, s, view(g), g->op_mask, words);
    }
    HIPC(hipGetLastError());
    g->epoch = 0;
    ++g->map_gen;
    return TSDF_OK;
}

tsdf_status tsdf_get_params(const tsdf_grid* g, tsdf_params* p)
{
    if (!g || !p)
        return fail(TSDF_BAD_ARG, "NULL grid or params");
    *p = g->p;
    return TSDF_OK;
}

tsdf_status tsdf_set_intrinsics(tsdf_grid* g, float fx, float fy, float cx, float cy)
{
    if (!g)
        return fail(TSDF_BAD_ARG, "NULL grid");
    if (!(std::isfinite(fx) && std::isfinite(fy) && std::isfinite(cx) && std::isfinite(cy)) || fx == 0.f ||
        fy == 0.f)
        return fail(TSDF_BAD_ARG, "intrinsics must be finite with fx, fy != 0");
    g->cam.fx = fx;
    g->cam.fy = fy;
    g->cam.ci = (float)g->p.height - cy;  // GPUVoxelGrid.cpp:259
    g->cam.cj = cx;                       // GPUVoxelGrid.cpp:260
    // getPoint's first factors per column and row (binary32 on the host, the
    // same IEEE division the device makes), for the depth-gather update
    {
        std::vector<float> cf((size_t)g->p.width), rf((size_t)g->p.height);
        for (int j = 0; j < g->p.width; ++j)
            cf[j] = ((float)j - g->cam.cj) / g->cam.fx;
        for (int i = 0; i < g->p.height; ++i)
            rf[i] = ((float)i - g->cam.ci) / g->cam.fy;
        DeviceScope ds(g->p.device);
        HIPC(hipMemcpy(g->colfac, cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(g->rowfac, rf.data(), rf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    g->intr_set = true;
    return TSDF_OK;
}

tsdf_status tsdf_integrate(tsdf_grid* g, const uint16_t* depth, int32_t pitch, const uint8_t* rgb,
                           const float E[16], int32_t sign, void* stream)
{
    if (!g || !E)
        return fail(TSDF_BAD_ARG, "NULL grid or pose");
    if (!g->intr_set)
        return fail(TSDF_NO_INTRINSICS, "tsdf_set_intrinsics was not called");
    tsdf_status cs = check_image_args(g, depth, pitch);
    if (cs != TSDF_OK)
        return cs;
    if (sign != 1 && sign != -1)
        return fail(TSDF_BAD_ARG, "sign must be +1 or -1");
    if (rgb && !g->p.with_color)
        return fail(TSDF_BAD_ARG, "colour given to a grid created without colour storage");
    if (!pose_ok(E))
        return fail(TSDF_BAD_POSE, "pose must be finite and rigid with last row 0 0 0 1");
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    OpDesc op;
    op.depth = depth;
    op.rgb = rgb;
    op.ptmap = g->ptmap;
    op.pitch = pitch;
    op.sign = sign;
    make_pose(E, &op.pose);
    const uint32_t epoch = ++g->epoch;
    ++g->map_gen;
    const int par = (int)(epoch & 1u);
    const dim3 cgrid((g->p.width + kTileW - 1) / kTileW, (g->p.height + kTileH - 1) / kTileH);
    {
        ProfScope ps(g, s, TSDF_K_COLLECT);
        hipLaunchKernelGGL(k_collect, cgrid, dim3(kThreads), 0, s, view(g), op, epoch, par);
    }
    const bool color = g->p.with_color != 0;
    {
        ProfScope ps(g, s, TSDF_K_INTEGRATE);
        if (g->p.storage == TSDF_STORAGE_F16)
            launch_integrate<__half>(g, op, par, s, color);
        else
            launch_integrate<float>(g, op, par, s, color);
    }
    HIPC(hipGetLastError());
    return TSDF_OK;
}

}  // extern "C"

namespace {

size_t stage_bytes(int n) { return offsetof(BatchStage, ops) + (size_t)n * sizeof(OpDesc); }

// The passes of a batch of n staged ops on stream s.
void enqueue_batch_passes(tsdf_grid* g, int n_ops, int per, hipStream_t s, bool profile)
{
    const bool color = g->p.with_color != 0;
    const uint32_t* ep = &g->stage_dev->epoch_base;
    for (int cb = 0, pass = 0; cb < n_ops; cb += per, ++pass) {
        const int n = (n_ops - cb) < per ? (n_ops - cb) : per;
        const dim3 cgrid((g->p.width + kBatchTileW - 1) / kBatchTileW, (g->p.height + kBatchTileH - 1) / kBatchTileH, n);
        const OpDesc* ops = g->stage_dev->ops + cb;
        {
            ProfScope ps(g, s, TSDF_K_COLLECT_BATCH, profile);
            hipLaunchKernelGGL(k_collect_batch, cgrid, dim3(kThreads), 0, s, view(g), ops, ep, pass, g->op_mask);
        }
        {
            ProfScope ps(g, s, TSDF_K_INTEGRATE_BATCH, profile);
            if (g->p.storage == TSDF_STORAGE_F16)
                launch_integrate_batch<__half>(g, ops, n, ep, pass, s, color);
            else
                launch_integrate_batch<float>(g, ops, n, ep, pass, s, color);
        }
    }
}

// Captures (staging copy + passes) for n ops in passes of `per` into one
// executable graph, unless the current one already has that shape.
tsdf_status ensure_batch_graph(tsdf_grid* g, int n_ops, int per)
{
    const int key[5] = {n_ops, per, persistent_grid(g), g->p.storage, g->p.with_color};
    if (g->batch_exec && memcmp(key, g->batch_key, sizeof(key)) == 0)
        return TSDF_OK;
    if (g->batch_exec) {
        HIPC(hipGraphExecDestroy(g->batch_exec));
        g->batch_exec = nullptr;
    }
    if (!g->cap_stream)
        HIPC(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    HIPC(hipStreamBeginCapture(g->cap_stream, hipStreamCaptureModeThreadLocal));
    (void)hipMemcpyAsync(g->stage_dev, g->stage_host, stage_bytes(n_ops), hipMemcpyHostToDevice, g->cap_stream);
    enqueue_batch_passes(g, n_ops, per, g->cap_stream, false);
    const hipError_t ce = hipStreamEndCapture(g->cap_stream, &graph);
    if (ce != hipSuccess || !graph)
        return fail(TSDF_HIP_ERROR, std::string("batch graph capture: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiateWithFlags(&g->batch_exec, graph, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        g->batch_exec = nullptr;
        return fail(TSDF_HIP_ERROR, std::string("batch graph instantiate: ") + hipGetErrorString(ie));
    }
    memcpy(g->batch_key, key, sizeof(key));
    return TSDF_OK;
}

}  // namespace

extern "C" {

tsdf_status tsdf_integrate_batch(tsdf_grid* g, int32_t n_ops, const uint16_t* const* depths, int32_t pitch,
                                 const uint8_t* const* rgbs, const float* E, const int32_t* signs, void* stream)
{
    if (!g || !depths || !E || !signs || n_ops < 0)
        return fail(TSDF_BAD_ARG, "NULL argument or negative op count");
    if (!g->intr_set)
        return fail(TSDF_NO_INTRINSICS, "tsdf_set_intrinsics was not called");
    if (pitch < g->p.width)
        return fail(TSDF_BAD_DIMS, "pitch smaller than the grid width");
    if (rgbs && !g->p.with_color)
        return fail(TSDF_BAD_ARG, "colour given to a grid created without colour storage");
    for (int k = 0; k < n_ops; ++k) {
        if (!depths[k])
            return fail(TSDF_BAD_ARG, "depth pointer of an op is NULL");
        if (signs[k] != 1 && signs[k] != -1)
            return fail(TSDF_BAD_ARG, "sign must be +1 or -1");
        if (!pose_ok(E + 16 * (size_t)k))
            return fail(TSDF_BAD_POSE, "pose of an op is not a finite rigid transform");
    }
    if (n_ops == 0)
        return TSDF_OK;
    DeviceScope ds(g->p.device);
    tsdf_status bs = ensure_batch_buffers(g, n_ops);
    if (bs != TSDF_OK)
        return bs;
    hipStream_t s = (hipStream_t)stream;
    const bool color = g->p.with_color != 0;
    const size_t frame = (size_t)g->p.height * g->p.width;
    // Ops go through the pinned staging buffer kMaxOps at a time; within that,
    // passes of `per` ops each run one collection and one brick update.  A
    // pass's point maps (per * H * W * 16 B) are rewritten by the next pass,
    // so a small pass keeps them in the last-level cache.
    const int per = g->tune.batch_ops_per_pass < g->batch_cap ? g->tune.batch_ops_per_pass : g->batch_cap;
    for (int sbase = 0; sbase < n_ops; sbase += kMaxOps) {
        const int sn = (n_ops - sbase) < kMaxOps ? (n_ops - sbase) : kMaxOps;
        // the staging buffer may still feed the previous copy
        if (g->staged)
            HIPC(hipEventSynchronize(g->staging_done));
        BatchStage* st = g->stage_host;
        for (int k = 0; k < sn; ++k) {
            OpDesc& op = st->ops[k];
            op.depth = depths[sbase + k];
            op.rgb = (color && rgbs) ? rgbs[sbase + k] : nullptr;
            op.ptmap = g->ptmap_batch + (size_t)(k % per) * frame;
            op.pitch = pitch;
            op.sign = signs[sbase + k];
            make_pose(E + 16 * (size_t)(sbase + k), &op.pose);
        }
        const int passes = (sn + per - 1) / per;
        st->epoch_base = g->epoch + 1;
        g->epoch += (uint32_t)passes;
        ++g->map_gen;
        // one captured graph (staging copy + all passes) when the shape
        // allows; per-kernel profiling needs the separate launches
        if (g->tune.batch_graph && !g->prof_on && n_ops <= kMaxOps) {
            tsdf_status r = ensure_batch_graph(g, sn, per);
            if (r != TSDF_OK)
                return r;
            HIPC(hipGraphLaunch(g->batch_exec, s));
        } else {
            HIPC(hipMemcpyAsync(g->stage_dev, st, stage_bytes(sn), hipMemcpyHostToDevice, s));
            enqueue_batch_passes(g, sn, per, s, true);
        }
        HIPC(hipEventRecord(g->staging_done, s));
        g->staged = true;
        HIPC(hipGetLastError());
    }
    return TSDF_OK;
}

tsdf_status tsdf_reserve_batch(tsdf_grid* g, int32_t n_ops)
{
    if (!g || n_ops < 0)
        return fail(TSDF_BAD_ARG, "NULL handle or negative op count");
    if (n_ops == 0)
        return TSDF_OK;
    DeviceScope ds(g->p.device);
    tsdf_status r = ensure_batch_buffers(g, n_ops);
    if (r != TSDF_OK || !g->tune.batch_graph || n_ops > kMaxOps)
        return r;
    // capture the graph of a batch of exactly n_ops now, so that such a
    // batch neither allocates nor instantiates
    const int per = g->tune.batch_ops_per_pass < g->batch_cap ? g->tune.batch_ops_per_pass : g->batch_cap;
    return ensure_batch_graph(g, n_ops, per);
}

tsdf_status tsdf_reintegrate_batch(tsdf_grid* g, int32_t n, const uint16_t* const* depths, int32_t pitch,
                                   const uint8_t* const* rgbs, const float* E_old, const float* E_new, void* stream)
{
    if (!g || !depths || !E_old || !E_new || n < 0)
        return fail(TSDF_BAD_ARG, "NULL argument or negative keyframe count");
    std::string err;
    const uint16_t** d = new const uint16_t*[2 * (size_t)n + 1];
    const uint8_t** c = rgbs ? new const uint8_t*[2 * (size_t)n + 1] : nullptr;
    float* E = new float[32 * (size_t)n + 16];
    int32_t* sg = new int32_t[2 * (size_t)n + 1];
    for (int k = 0; k < n; ++k) {
        d[2 * k] = d[2 * k + 1] = depths[k];
        if (c)
            c[2 * k] = c[2 * k + 1] = rgbs[k];
        memcpy(E + 32 * (size_t)k, E_old + 16 * (size_t)k, 16 * sizeof(float));
        memcpy(E + 32 * (size_t)k + 16, E_new + 16 * (size_t)k, 16 * sizeof(float));
        sg[2 * k] = -1;
        sg[2 * k + 1] = 1;
    }
    const tsdf_status r = tsdf_integrate_batch(g, 2 * n, d, pitch, c, E, sg, stream);
    delete[] d;
    delete[] c;
    delete[] E;
    delete[] sg;
    return r;
}

tsdf_status tsdf_raycast(tsdf_grid* g, const float E[16], float* points, float* colors, float* depth,
                         uint8_t* ub_flags, void* stream)
{
    if (!g || !E || !points)
        return fail(TSDF_BAD_ARG, "NULL grid, pose or points");
    if (!g->intr_set)
        return fail(TSDF_NO_INTRINSICS, "tsdf_set_intrinsics was not called");
    if (!pose_ok(E))
        return fail(TSDF_BAD_POSE, "pose must be finite and rigid with last row 0 0 0 1");
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    Pose P;
    make_pose(E, &P);
    Footprint fp;
    memset(&fp, 0, sizeof(fp));
    {
        ProfScope ps(g, s, TSDF_K_RAYCAST);
        launch_raycast<false>(g, P, points, colors, depth, ub_flags, fp, s);
    }
    HIPC(hipGetLastError());
    return TSDF_OK;
}

tsdf_status tsdf_raycast_footprint(tsdf_grid* g, const float E[16], uint64_t* distinct_voxels,
                                   uint64_t* distinct_keys, void* stream)
{
    if (!g || !E || !distinct_voxels || !distinct_keys)
        return fail(TSDF_BAD_ARG, "NULL argument");
    if (!g->intr_set)
        return fail(TSDF_NO_INTRINSICS, "tsdf_set_intrinsics was not called");
    if (!pose_ok(E))
        return fail(TSDF_BAD_POSE, "pose must be finite and rigid with last row 0 0 0 1");
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    const size_t words = (size_t)g->p.n_bricks * 16;
    if (!g->fp_bits) {
        g->fp_key_mask = (1u << 23) - 1u;
        HIPC(hipMalloc(&g->fp_bits, words * sizeof(uint32_t)));
        HIPC(hipMalloc(&g->fp_keys, ((size_t)g->fp_key_mask + 1) * sizeof(unsigned long long)));
        HIPC(hipMalloc(&g->fp_counts, 4 * sizeof(unsigned long long)));
        HIPC(hipMalloc(&g->fp_points, (size_t)g->p.height * g->p.width * 3 * sizeof(float)));
    }
    HIPC(hipMemsetAsync(g->fp_bits, 0, words * sizeof(uint32_t), s));
    HIPC(hipMemsetAsync(g->fp_keys, 0xFF, ((size_t)g->fp_key_mask + 1) * sizeof(unsigned long long), s));
    HIPC(hipMemsetAsync(g->fp_counts, 0, 4 * sizeof(unsigned long long), s));
    Pose P;
    make_pose(E, &P);
    Footprint fp;
    fp.vox_bits = g->fp_bits;
    fp.keys = g->fp_keys;
    fp.key_mask = g->fp_key_mask;
    fp.counts = g->fp_counts;
    launch_raycast<true>(g, P, g->fp_points, nullptr, nullptr, nullptr, fp, s);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_popcount, dim3(g->n_cu * 4), dim3(kThreads), 0, s, g->fp_bits, words, g->fp_counts + 2);
    HIPC(hipGetLastError());
    unsigned long long c[4];
    HIPC(hipMemcpyAsync(c, g->fp_counts, sizeof(c), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (c[1])
        return fail(TSDF_MAP_FULL, "footprint key set overflowed: " + std::to_string(c[0]) + " keys inserted, " +
                                       std::to_string(c[1]) + " refused");
    *distinct_voxels = c[2];
    *distinct_keys = c[0];
    return TSDF_OK;
}

void tsdf_default_tuning(tsdf_tuning* t)
{
    if (!t)
        return;
    t->raycast_xcd_map = 0;
    t->raycast_skip = 2;
    t->raycast_budget = 12;
    t->integrate_wg_per_cu = 6;
    t->batch_ops_per_pass = 8;
    t->batch_graph = 1;
    t->raycast_order = 2;
}

tsdf_status tsdf_set_tuning(tsdf_grid* g, const tsdf_tuning* t)
{
    if (!g || !t)
        return fail(TSDF_BAD_ARG, "NULL grid or tuning");
    if (t->raycast_xcd_map < 0 || t->raycast_xcd_map > 1 || t->raycast_skip < 0 || t->raycast_skip > 2 ||
        t->raycast_budget < 0 || t->raycast_budget > 4096 || t->integrate_wg_per_cu < 1 ||
        t->integrate_wg_per_cu > 32 || t->batch_ops_per_pass < 1 || t->batch_ops_per_pass > kMaxOps ||
        t->batch_graph < 0 || t->batch_graph > 1 || t->raycast_order < 0 || t->raycast_order > 2)
        return fail(TSDF_BAD_ARG, "tuning field out of range (workgroups per CU in [1, 32], raycast order and "
                                  "empty-space skip 0, 1 or 2, flags 0 or 1)");
    g->tune = *t;
    return TSDF_OK;
}

tsdf_status tsdf_get_tuning(const tsdf_grid* g, tsdf_tuning* t)
{
    if (!g || !t)
        return fail(TSDF_BAD_ARG, "NULL grid or tuning");
    *t = g->tune;
    return TSDF_OK;
}

tsdf_status tsdf_profile(tsdf_grid* g, int32_t enable)
{
    if (!g)
        return fail(TSDF_BAD_ARG, "NULL grid");
    g->prof_on = enable != 0;
    return TSDF_OK;
}

tsdf_status tsdf_profile_read(tsdf_grid* g, tsdf_kernel_times* out, int32_t reset)
{
    if (!g || !out)
        return fail(TSDF_BAD_ARG, "NULL grid or output");
    DeviceScope ds(g->p.device);
    tsdf_status st = prof_flush(g);
    if (st != TSDF_OK)
        return st;
    for (int k = 0; k < TSDF_N_KERNEL_KINDS; ++k) {
        out->ms[k] = g->prof_ms[k];
        out->launches[k] = g->prof_n[k];
        if (reset) {
            g->prof_ms[k] = 0.0;
            g->prof_n[k] = 0;
        }
    }
    return TSDF_OK;
}

tsdf_status tsdf_stats(tsdf_grid* g, tsdf_stats_out* out, void* stream)
{
    if (!g || !out)
        return fail(TSDF_BAD_ARG, "NULL grid or output");
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    HIPC(hipMemcpyAsync(g->ctrl_host, g->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(g->stat_host, g->stat, kStatWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    const Ctrl& c = *g->ctrl_host;
    unsigned long long tot[kStCount];
    for (int k = 0; k < kStCount; ++k) {
        tot[k] = 0;
        for (int l = 0; l < kStatLanes; ++l)
            tot[k] += g->stat_host[((size_t)k * kStatLanes + l) * kStatStride];
    }
    out->n_bricks = c.alloc_top < g->p.n_bricks ? c.alloc_top : g->p.n_bricks;
    out->capacity = g->p.n_bricks;
    out->overflow = c.overflow;
    out->last_touched = g->epoch ? c.touched_count[g->epoch & 1u] : 0;
    out->total_touched = c.total_touched;
    out->total_new = tot[kStNew];
    out->total_updated = tot[kStUpdated];
    out->total_valid_px = tot[kStValid];
    out->ray_hits = tot[kStRayHits];
    out->ray_ub = tot[kStRayUb];
    out->ray_samples = tot[kStRaySamples];
    out->ray_probes = tot[kStRayProbes];
    return TSDF_OK;
}

tsdf_status tsdf_sync(tsdf_grid* g, void* stream)
{
    tsdf_stats_out st;
    tsdf_status r = tsdf_stats(g, &st, stream);
    if (r != TSDF_OK)
        return r;
    if (st.overflow)
        return fail(TSDF_MAP_FULL, "brick pool or hash table overflowed; " + std::to_string(st.overflow) +
                                       " allocations refused since the last reset");
    return TSDF_OK;
}

tsdf_status tsdf_export_bricks(tsdf_grid* g, int32_t* coords, float* sdf, float* w, float* rgb, uint64_t cap,
                               uint64_t* n_out, void* stream)
{
    if (!g || !n_out)
        return fail(TSDF_BAD_ARG, "NULL grid or count output");
    tsdf_stats_out st;
    tsdf_status r = tsdf_stats(g, &st, stream);
    if (r != TSDF_OK)
        return r;
    *n_out = st.n_bricks;
    const uint64_t n = st.n_bricks < cap ? st.n_bricks : cap;
    if (n == 0)
        return TSDF_OK;
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    const int grid = (int)(n < 65536 ? n : 65536);
    float* c = g->p.with_color ? rgb : nullptr;
    if (g->p.storage == TSDF_STORAGE_F16)
        hipLaunchKernelGGL(k_export<__half>, dim3(grid), dim3(kThreads), 0, s, view(g), (unsigned)n, coords, sdf, w, c);
    else
        hipLaunchKernelGGL(k_export<float>, dim3(grid), dim3(kThreads), 0, s, view(g), (unsigned)n, coords, sdf, w, c);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s));
    return TSDF_OK;
}

}  // extern "C"

#include "tsdf_render.hpp"
#include "tsdf_maintain.hpp"
#include "tsdf_mc.hpp"
#include "tsdf_icp.hpp"
#include "tsdf_sched.hpp"
