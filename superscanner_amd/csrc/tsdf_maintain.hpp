// tsdf_maintain.hpp -- map maintenance and surface extraction: erase empty
// bricks (cleanMemory) and a triangle mesh of the zero level set.  Included
// once, at the end of tsdf_hip.hip.
//
// cleanMemory / cleanTag, GPUVoxelGrid.cpp:185-196, 785-805: every brick whose
// 512 weights are all zero leaves the map.  Here the surviving bricks are
// compacted to ids 0..n-1 (ids are not observable) and the table is rebuilt.
//
// marchingCubesHost, GPUVoxelGrid.cpp:200-251, copies the bricks into an
// Open3D ScalableTSDFVolume and calls its ExtractTriangleMesh on the host.
// Here the mesh is extracted on the GPU by marching tetrahedra: every cell of
// 8 voxel centres (x..x+1, y..y+1, z..z+1), all observed (w > 0), is split
// into 6 tetrahedra around its diagonal from corner 0 to corner 7 (corner
// c = (c & 1, c >> 1 & 1, c >> 2 & 1)); the same split in every cell makes
// neighbouring cells agree on their shared faces, so the surface is closed
// wherever the field is observed.  Vertices lie on tetrahedron edges whose
// ends change sign, at p = pa + t (pb - pa), t = sa / (sa - sb) with a the
// inside (sdf < 0) corner, so neighbours compute shared vertices bit for bit;
// each triangle faces the positive (free-space) side.
#pragma once

namespace {

// ---------------------------------------------------------------------------
// clean
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_alive(GridView g, unsigned int n, uint8_t* alive)
{
    const int lane = threadIdx.x & 63;
    const unsigned int nw = gridDim.x * (kThreads / 64);
    for (unsigned int b = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); b < n; b += nw) {
        const T* sd = (const T*)g.payload + (size_t)b * (2 * kVPB);
        const bool fresh = (g.coords[b].w & 1) != 0;  // allocated, never written: all zero
        bool any = false;
        for (int v = lane; v < kVPB && !fresh; v += 64) {
            float s, w;
            ld_pair(sd, v, s, w);
            any = any || (w != 0.0f);
        }
        const bool r = __ballot(any) != 0ull;
        if (lane == 0)
            alive[b] = r ? 1 : 0;
    }
}

// new id -> old id: moves coords, payload (and colour) of the kept bricks into
// the compact staging buffers
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather_bricks(GridView g, const uint32_t* old_of, unsigned int n,
                                                            int4* c_out, T* p_out, float* rgb_out)
{
    for (unsigned int j = blockIdx.x; j < n; j += gridDim.x) {
        const uint32_t i = old_of[j];
        if (threadIdx.x == 0)
            c_out[j] = g.coords[i];
        const T* src = (const T*)g.payload + (size_t)i * (2 * kVPB);
        T* dst = p_out + (size_t)j * (2 * kVPB);
        for (int k = threadIdx.x; k < 2 * kVPB; k += kThreads)
            dst[k] = src[k];
        if (rgb_out)
            for (int k = threadIdx.x; k < 3 * kVPB; k += kThreads)
                rgb_out[(size_t)j * 3 * kVPB + k] = g.rgb[(size_t)i * 3 * kVPB + k];
    }
}

__global__ __launch_bounds__(kThreads) void k_rehash_clear(GridView g)
{
    const size_t n = (size_t)g.mask + 1;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        g.slots[i] = kSlotEmpty;
        g.slot_epoch[i] = 0u;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < kSbWords; i += stride)
        g.sb_bits[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int a = 0; a < 3; ++a) {
            g.ctrl->bbox_min[a] = 0x7FFFFFFF;
            g.ctrl->bbox_max[a] = (int)0x80000000u;
        }
        g.ctrl->sb_version += 1u;  // the bits were cleared
    }
}

// inserts brick j (already at its final id) under its reference key
__global__ __launch_bounds__(kThreads) void k_rehash_insert(GridView g, unsigned int n)
{
    for (unsigned int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const int4 c = g.coords[j];
        const uint32_t key = prime_hash(c.x, c.y, c.z);
        const unsigned long long word = ((unsigned long long)key << 32) | j;
        uint32_t b = mix32(key) & g.mask;
        for (uint32_t p = 0; p <= g.mask; ++p) {
            if (atomicCAS(&g.slots[b], kSlotEmpty, word) == kSlotEmpty)
                break;
            b = (b + 1) & g.mask;
        }
        const int sb = sb_index(c.x, c.y, c.z);
        if (sb >= 0 && !(atomicOr(&g.sb_bits[sb >> 5], 1u << (sb & 31)) & (1u << (sb & 31))))
            atomicAdd(&g.ctrl->sb_version, 1u);
        grow_bbox(g.ctrl, c.x, c.y, c.z);
        if (j == 0)
            g.ctrl->alloc_top = n;
    }
}

__global__ void k_set_alloc_top(Ctrl* ctrl, unsigned int n) { ctrl->alloc_top = n; }

__global__ void k_fill_u64(unsigned long long* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------------

// The 6 tetrahedra of a cell, corners by index c = x + 2y + 4z.
__constant__ int kTets[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};

// A triangle as the three tetrahedron edges its vertices lie on: vertex q is
// on the edge from inside corner a[q] to outside corner b[q].
struct TriEdges {
    int8_t a[3], b[3];
};

__device__ __forceinline__ void edge_point(const float (&p)[8][3], const float (&s)[8], int a, int b, float* out)
{
    const float t = s[a] / (s[a] - s[b]);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        out[k] = p[a][k] + t * (p[b][k] - p[a][k]);
}

// Triangles of tetrahedron `tet` of a cell; returns how many (0..2).
__device__ __forceinline__ int tet_triangles(const float (&p)[8][3], const float (&s)[8], int tet, TriEdges* out)
{
    const int c[4] = {kTets[tet][0], kTets[tet][1], kTets[tet][2], kTets[tet][3]};
    int in[4], out_[4], ni = 0, no = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (s[c[k]] < 0.0f)
            in[ni++] = c[k];
        else
            out_[no++] = c[k];
    }
    if (ni == 0 || no == 0)
        return 0;
    int nt;
    if (ni == 1) {
        out[0] = TriEdges{{(int8_t)in[0], (int8_t)in[0], (int8_t)in[0]},
                          {(int8_t)out_[0], (int8_t)out_[1], (int8_t)out_[2]}};
        nt = 1;
    } else if (no == 1) {
        out[0] = TriEdges{{(int8_t)in[0], (int8_t)in[1], (int8_t)in[2]},
                          {(int8_t)out_[0], (int8_t)out_[0], (int8_t)out_[0]}};
        nt = 1;
    } else {
        // the quad e00 e01 e11 e10 (e_ij: inside i to outside j) as two triangles
        out[0] = TriEdges{{(int8_t)in[0], (int8_t)in[0], (int8_t)in[1]},
                          {(int8_t)out_[0], (int8_t)out_[1], (int8_t)out_[1]}};
        out[1] = TriEdges{{(int8_t)in[0], (int8_t)in[1], (int8_t)in[1]},
                          {(int8_t)out_[0], (int8_t)out_[1], (int8_t)out_[0]}};
        nt = 2;
    }
    // face the free-space side: from the inside corners toward the outside
    float d[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < ni; ++k)
        for (int q = 0; q < 3; ++q)
            d[q] -= p[in[k]][q] / (float)ni;
    for (int k = 0; k < no; ++k)
        for (int q = 0; q < 3; ++q)
            d[q] += p[out_[k]][q] / (float)no;
    for (int t = 0; t < nt; ++t) {
        float v0[3], v1[3], v2[3];
        edge_point(p, s, out[t].a[0], out[t].b[0], v0);
        edge_point(p, s, out[t].a[1], out[t].b[1], v1);
        edge_point(p, s, out[t].a[2], out[t].b[2], v2);
        const float a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const float b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        const float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] < 0.0f) {
            const int8_t ta = out[t].a[1], tb = out[t].b[1];
            out[t].a[1] = out[t].a[2];
            out[t].b[1] = out[t].b[2];
            out[t].a[2] = ta;
            out[t].b[2] = tb;
        }
    }
    return nt;
}

// Key of the vertex on the edge from voxel (gx, gy, gz) (inside) to its
// neighbour at offset (dx, dy, dz) (outside): the voxel's global index,
// biased by 2^18 per axis (19 bits each), and the offset (5 bits).  Equal
// keys are equal vertices bit for bit.
constexpr int kKeyBias = 1 << 18;

__device__ __forceinline__ unsigned long long vertex_key(int gx, int gy, int gz, int dx, int dy, int dz)
{
    return ((unsigned long long)(gx + kKeyBias) << 43) | ((unsigned long long)(gy + kKeyBias) << 24) |
           ((unsigned long long)(gz + kKeyBias) << 5) | (unsigned long long)((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1));
}

// One workgroup per brick, two cells per thread.  tris: [cap][3][3] world
// coordinates; cols (kColor): [cap][3][3] vertex colours interpolated like
// the positions; keys (or null): [cap][3] vertex keys; count: total
// triangles (may exceed cap).  A brick that was allocated but never written
// (coords.w bit 0) counts as unobserved.
template <typename T, bool kColor>
__global__ __launch_bounds__(kThreads) void k_mesh(GridView g, unsigned int n, float* tris, float* cols,
                                                   unsigned long long* keys, unsigned long long cap,
                                                   unsigned long long* count)
{
    __shared__ int nb[8];  // brick ids of (dx, dy, dz) = (c & 1, c >> 1 & 1, c >> 2 & 1)
    __shared__ int4 bc;
    for (unsigned int b = blockIdx.x; b < n; b += gridDim.x) {
        __syncthreads();
        if (threadIdx.x == 0)
            bc = g.coords[b];
        __syncthreads();
        if (threadIdx.x < 8) {
            const int c = threadIdx.x;
            int id = c == 0 ? (int)b : find_brick(g, bc.x + (c & 1), bc.y + ((c >> 1) & 1), bc.z + ((c >> 2) & 1));
            if (id >= 0 && (g.coords[id].w & 1))
                id = -1;
            nb[c] = id;
        }
        __syncthreads();
        for (int cell = threadIdx.x; cell < kVPB; cell += kThreads) {
            const int x = cell & 7, y = (cell >> 3) & 7, z = cell >> 6;
            float s[8], p[8][3];
            int vid[8], vbrick[8];
            bool ok = true;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int vx = x + (c & 1), vy = y + ((c >> 1) & 1), vz = z + ((c >> 2) & 1);
                const int which = (vx >> 3) | ((vy >> 3) << 1) | ((vz >> 3) << 2);
                const int id = nb[which];
                const int v = (vx & 7) + 8 * (vy & 7) + 64 * (vz & 7);
                float w = 0.f, sv = 0.f;
                if (id >= 0)
                    ld_pair((const T*)g.payload + (size_t)id * (2 * kVPB), v, sv, w);
                ok = ok && id >= 0 && w > 0.0f;
                s[c] = sv;
                vid[c] = v;
                vbrick[c] = id;
                p[c][0] = voxel_mid(vx & 7, bc.x + (vx >> 3), g.vbl, g.vl);
                p[c][1] = voxel_mid(vy & 7, bc.y + (vy >> 3), g.vbl, g.vl);
                p[c][2] = voxel_mid(vz & 7, bc.z + (vz >> 3), g.vbl, g.vl);
            }
            TriEdges tri[12];
            int nt = 0;
            if (ok)
                for (int t = 0; t < 6; ++t)
                    nt += tet_triangles(p, s, t, tri + nt);
            // wave-aggregated slots in the output
            int incl = nt;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if ((int)__lane_id() >= o)
                    incl += v;
            }
            const int total = __shfl(incl, 63);
            unsigned long long base = 0;
            if (__lane_id() == 63 && total)
                base = atomicAdd(count, (unsigned long long)total);
            base = __shfl(base, 63);
            const unsigned long long first = base + (unsigned long long)(incl - nt);
            for (int t = 0; t < nt; ++t) {
                const unsigned long long k = first + t;
                if (k >= cap)
                    break;
                for (int q = 0; q < 3; ++q) {
                    const int a = tri[t].a[q], bb = tri[t].b[q];
                    float v[3];
                    edge_point(p, s, a, bb, v);
                    for (int r = 0; r < 3; ++r)
                        tris[k * 9 + q * 3 + r] = v[r];
                    if (keys) {
                        const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
                        keys[k * 3 + q] = vertex_key(bc.x * 8 + x + ax, bc.y * 8 + y + ay, bc.z * 8 + z + az,
                                                     (bb & 1) - ax, ((bb >> 1) & 1) - ay, ((bb >> 2) & 1) - az);
                    }
                    if (kColor) {
                        const float u = s[a] / (s[a] - s[bb]);
                        const float* ca = g.rgb + (size_t)vbrick[a] * 3 * kVPB + vid[a];
                        const float* cb = g.rgb + (size_t)vbrick[bb] * 3 * kVPB + vid[bb];
                        for (int r = 0; r < 3; ++r) {
                            const float x0 = ca[r * kVPB], x1 = cb[r * kVPB];
                            cols[k * 9 + q * 3 + r] = x0 + u * (x1 - x0);
                        }
                    }
                }
            }
        }
    }
}

}  // namespace

extern "C" {

tsdf_status tsdf_clean(tsdf_grid* g, uint64_t* n_removed, void* stream)
{
    if (!g)
        return fail(TSDF_BAD_ARG, "NULL grid");
    tsdf_stats_out st;
    tsdf_status r = tsdf_stats(g, &st, stream);
    if (r != TSDF_OK)
        return r;
    const unsigned int n = (unsigned int)st.n_bricks;
    if (n_removed)
        *n_removed = 0;
    if (n == 0)
        return TSDF_OK;
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    ++g->map_gen;  // the table and superblock bits are rebuilt
    const bool f16 = g->p.storage == TSDF_STORAGE_F16;
    const size_t tsz = f16 ? sizeof(__half) : sizeof(float);
    uint8_t* alive_d = nullptr;
    HIPC(hipMalloc(&alive_d, n));
    const int grid = g->n_cu * 8;
    if (f16)
        hipLaunchKernelGGL(k_alive<__half>, dim3(grid), dim3(kThreads), 0, s, view(g), n, alive_d);
    else
        hipLaunchKernelGGL(k_alive<float>, dim3(grid), dim3(kThreads), 0, s, view(g), n, alive_d);
    std::string err;
    uint8_t* alive = (uint8_t*)malloc(n);
    uint32_t* old_of = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* old_of_d = nullptr;
    int4* c_tmp = nullptr;
    void* p_tmp = nullptr;
    float* rgb_tmp = nullptr;
    unsigned int kept = 0;
    tsdf_status out = TSDF_OK;
    auto check = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && out == TSDF_OK)
            out = fail(TSDF_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
        return out == TSDF_OK;
    };
    if (!alive || !old_of) {
        out = fail(TSDF_HIP_ERROR, "host allocation failed");
    } else if (check(hipMemcpyAsync(alive, alive_d, n, hipMemcpyDeviceToHost, s), "copy flags") &&
               check(hipStreamSynchronize(s), "synchronize")) {
        for (unsigned int i = 0; i < n; ++i)
            if (alive[i])
                old_of[kept++] = i;
        if (kept < n && kept > 0 && check(hipMalloc(&old_of_d, (size_t)kept * sizeof(uint32_t)), "hipMalloc") &&
            check(hipMalloc(&c_tmp, (size_t)kept * sizeof(int4)), "hipMalloc") &&
            check(hipMalloc(&p_tmp, (size_t)kept * 2 * kVPB * tsz), "hipMalloc") &&
            (!g->rgb || check(hipMalloc(&rgb_tmp, (size_t)kept * 3 * kVPB * sizeof(float)), "hipMalloc")) &&
            check(hipMemcpyAsync(old_of_d, old_of, (size_t)kept * sizeof(uint32_t), hipMemcpyHostToDevice, s),
                  "copy ids")) {
            const int gg = (int)(kept < 65536u ? kept : 65536u);
            if (f16)
                hipLaunchKernelGGL(k_gather_bricks<__half>, dim3(gg), dim3(kThreads), 0, s, view(g), old_of_d, kept,
                                   c_tmp, (__half*)p_tmp, rgb_tmp);
            else
                hipLaunchKernelGGL(k_gather_bricks<float>, dim3(gg), dim3(kThreads), 0, s, view(g), old_of_d, kept,
                                   c_tmp, (float*)p_tmp, rgb_tmp);
            if (check(hipMemcpyAsync(g->coords, c_tmp, (size_t)kept * sizeof(int4), hipMemcpyDeviceToDevice, s),
                      "copy coords") &&
                check(hipMemcpyAsync(g->payload, p_tmp, (size_t)kept * 2 * kVPB * tsz, hipMemcpyDeviceToDevice, s),
                      "copy payload") &&
                (!rgb_tmp || check(hipMemcpyAsync(g->rgb, rgb_tmp, (size_t)kept * 3 * kVPB * sizeof(float),
                                                  hipMemcpyDeviceToDevice, s),
                                   "copy colour"))) {
                hipLaunchKernelGGL(k_rehash_clear, dim3(g->n_cu * 4), dim3(kThreads), 0, s, view(g));
                hipLaunchKernelGGL(k_rehash_insert, dim3(g->n_cu * 4), dim3(kThreads), 0, s, view(g), kept);
                check(hipGetLastError(), "rehash");
            }
        } else if (kept == 0 && out == TSDF_OK) {
            hipLaunchKernelGGL(k_rehash_clear, dim3(g->n_cu * 4), dim3(kThreads), 0, s, view(g));
            hipLaunchKernelGGL(k_set_alloc_top, dim3(1), dim3(1), 0, s, g->ctrl, 0u);
            check(hipGetLastError(), "rehash");
        }
        check(hipStreamSynchronize(s), "synchronize");
    }
    (void)hipFree(alive_d);
    (void)hipFree(old_of_d);
    (void)hipFree(c_tmp);
    (void)hipFree(p_tmp);
    (void)hipFree(rgb_tmp);
    free(alive);
    free(old_of);
    if (out == TSDF_OK && n_removed)
        *n_removed = n - kept;
    return out;
}

tsdf_status tsdf_extract_mesh(tsdf_grid* g, float* triangles, float* colors, uint64_t* vertex_keys, uint64_t cap,
                              uint64_t* n_triangles, void* stream)
{
    if (!g || !n_triangles || (cap && !triangles))
        return fail(TSDF_BAD_ARG, "NULL grid, count or triangle buffer");
    if (colors && !g->rgb)
        return fail(TSDF_BAD_ARG, "vertex colours asked of a grid without colour");
    tsdf_stats_out st;
    tsdf_status r = tsdf_stats(g, &st, stream);
    if (r != TSDF_OK)
        return r;
    *n_triangles = 0;
    const unsigned int n = (unsigned int)st.n_bricks;
    if (n == 0)
        return TSDF_OK;
    if (vertex_keys) {
        // every block ever inserted lies in the control block's bounding box
        const Ctrl& c = *g->ctrl_host;
        for (int a = 0; a < 3; ++a)
            if (c.bbox_min[a] < -(kKeyBias / 8) || c.bbox_max[a] >= kKeyBias / 8 - 1)
                return fail(TSDF_BAD_ARG, "map extent beyond the vertex key range (+-32767 blocks per axis)");
    }
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* cnt = nullptr;
    HIPC(hipMalloc(&cnt, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, s, cnt, (size_t)1);
    unsigned long long* keys = cap > 0 ? (unsigned long long*)vertex_keys : nullptr;
    const int grid = (int)(n < 65536u ? n : 65536u);
    const bool f16 = g->p.storage == TSDF_STORAGE_F16;
    const bool col = colors != nullptr && cap > 0;
    const unsigned long long c = (unsigned long long)cap;
    if (f16 && col)
        hipLaunchKernelGGL((k_mesh<__half, true>), dim3(grid), dim3(kThreads), 0, s, view(g), n, triangles, colors, keys,
                           c, cnt);
    else if (f16)
        hipLaunchKernelGGL((k_mesh<__half, false>), dim3(grid), dim3(kThreads), 0, s, view(g), n, triangles, colors,
                           keys, c, cnt);
    else if (col)
        hipLaunchKernelGGL((k_mesh<float, true>), dim3(grid), dim3(kThreads), 0, s, view(g), n, triangles, colors, keys,
                           c, cnt);
    else
        hipLaunchKernelGGL((k_mesh<float, false>), dim3(grid), dim3(kThreads), 0, s, view(g), n, triangles, colors,
                           keys, c, cnt);
    unsigned long long h = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    (void)hipFree(cnt);
    if (e != hipSuccess)
        return fail(TSDF_HIP_ERROR, std::string("mesh extraction: ") + hipGetErrorString(e));
    *n_triangles = h;
    return TSDF_OK;
}

}  // extern "C"
