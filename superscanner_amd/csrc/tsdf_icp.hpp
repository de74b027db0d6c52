// tsdf_icp.hpp -- point maps, normal maps and projective ICP between two depth
// frames.  Included once, at the end of tsdf_hip.hip.
//
//   CreatePcdFromDepthView        imageutil.cpp:89-111     k_depth_to_points
//   CalculateNormalsFromPCDView   imageutil.cpp:116-157    k_normals
//   "Closest Point Kernel"        kokkosICP.cpp:69-115     k_icp_associate
//   host sums of an iteration     kokkosICP.cpp:117-171    k_icp_moments
//   kokkosICP (depth views)       kokkosICP.cpp:23-191     tsdf_icp
//
// Point and normal maps are row-major [H][W][3] float (the reference's Views
// are LayoutLeft Kvec3f [H][W]; the layout is internal to it).  The per-pixel
// kernels follow the reference's float operation order, so their outputs are
// bit-identical to it.  The iteration's sums are taken in double by a
// deterministic two-level reduction (fixed grid, fixed tree), where the
// reference sums on the host in pixel order (and the RMSE in float).
#pragma once

#include <mutex>

// Scratch of tsdf_icp_moments_noalloc: the per-workgroup partial sums and the
// final sums of k_icp_moments, (512 + 1) x 18 doubles = 73 872 bytes.  A
// static device array has one copy per device, each at its own address, held
// until the process ends.  On the host its name is only a stand-in: kernels
// and copies get the address hipGetSymbolAddress returns for the current
// device (icp_scratch_ptr).
static __device__ __align__(256) double scratch[(512 + 1) * TSDF_ICP_MOMENTS];

namespace {

constexpr int kIcpMoments = TSDF_ICP_MOMENTS;
constexpr int kIcpBlocks = 512;  // partial sums of k_icp_moments
static_assert(sizeof(scratch) == (size_t)(kIcpBlocks + 1) * kIcpMoments * sizeof(double),
              "the no-allocation scratch holds the partial and final sums");

__device__ __forceinline__ f3 ld3(const float* p, unsigned int k)
{
    return f3{p[3 * k], p[3 * k + 1], p[3 * k + 2]};
}

__device__ __forceinline__ void st3(float* p, unsigned int k, f3 v)
{
    p[3 * k] = v.x;
    p[3 * k + 1] = v.y;
    p[3 * k + 2] = v.z;
}

__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// Kvec3f::cross, kokkosFunctions.h:240-247
__device__ __forceinline__ f3 cross3(f3 a, f3 b)
{
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Kvec3f::length, kokkosFunctions.h:169-171
__device__ __forceinline__ float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }

// d = (float)depth / 1000 and getPoint's back-projection, one pixel a thread
__global__ __launch_bounds__(kThreads) void k_depth_to_points(const uint16_t* depth, int H, int W, Camera cam,
                                                              float* pts)
{
    const unsigned int n = (unsigned int)H * (unsigned int)W;
    for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = (int)(k / (unsigned int)W), j = (int)(k % (unsigned int)W);
        const float d = (float)depth[k] / 1000.0f;
        st3(pts, k, get_point(cam, i, j, d));
    }
}

// (0, 0, 1) on the border and where a 4-neighbour has z == 0; otherwise the
// normalised sum of the four neighbour cross products.
__global__ __launch_bounds__(kThreads) void k_normals(const float* pts, int H, int W, float* nrm)
{
    const unsigned int n = (unsigned int)H * (unsigned int)W;
    for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = (int)(k / (unsigned int)W), j = (int)(k % (unsigned int)W);
        f3 out{0.0f, 0.0f, 1.0f};
        if (i >= 1 && i < H - 1 && j >= 1 && j < W - 1) {
            const f3 p = ld3(pts, k);
            const f3 pd = ld3(pts, k + W), pu = ld3(pts, k - W);  // i + 1, i - 1
            const f3 pr = ld3(pts, k + 1), pl = ld3(pts, k - 1);  // j + 1, j - 1
            if (p.z != 0.0f && pd.z != 0.0f && pu.z != 0.0f && pr.z != 0.0f && pl.z != 0.0f) {
                const f3 n1 = cross3(sub3(pd, p), sub3(pr, p));
                const f3 n2 = cross3(sub3(pl, p), sub3(pd, p));
                const f3 n3 = cross3(sub3(pr, p), sub3(pu, p));
                const f3 n4 = cross3(sub3(pu, p), sub3(pl, p));
                f3 s{n1.x + n2.x, n1.y + n2.y, n1.z + n2.z};
                s = f3{s.x + n3.x, s.y + n3.y, s.z + n3.z};
                s = f3{s.x + n4.x, s.y + n4.y, s.z + n4.z};
                const float l = len3(s);
                out = f3{s.x / l, s.y / l, s.z / l};
            }
        }
        st3(nrm, k, out);
    }
}

struct IcpArgs {
    const float* x;
    const float* y;
    const float* yn;
    float T[12];  // Txy, 3x4 row-major
    Camera cam;
    int H, W;
    int n_search;
    float max_dist;
    int type;
    float* out;
};

// For every pixel of x with z != 0: px = Txy x, projected with the reference
// camera; the candidates y(k, l) of the window k in [ci - n, ci + n), l in
// [cj - n, cj + n) are scanned row by row, and the first strictly smallest
// distance below max_dist wins (point-to-point: |y - px|, out = y;
// point-to-plane: |(y - px) . n_y| over y with n_y.x != 0, out = px + d n_y).
// (0, 0, 0) where nothing qualifies.
//
// One workgroup per 16 x 16 pixels of x.  The union of its pixels' windows is
// staged in LDS (y, and the normals for point to plane) when it holds at most
// kIcpLds candidates; the scan then reads LDS, in the same order, else global
// memory.  kIcpLds = 0 always scans global memory.
#ifndef TSDF_ICP_LDS
#define TSDF_ICP_LDS 1024
#endif
constexpr int kIcpTile = 16;
constexpr int kIcpLds = TSDF_ICP_LDS;  // staged candidates per workgroup
constexpr int kIcpLdsAlloc = kIcpLds > 0 ? kIcpLds : 1;

// candidate (r, c): from the staged box (origin r0, c0, width bw) or global memory
template <bool kLds>
__device__ __forceinline__ f3 icp_cand(const float* lx, const float* ly, const float* lz, int r, int c, int r0,
                                      int c0, int bw, const float* g, unsigned int q)
{
    if (kLds) {
        const int o = (r - r0) * bw + (c - c0);
        return f3{lx[o], ly[o], lz[o]};
    }
    return ld3(g, q);
}

template <bool kLds, bool kPlane>
__device__ __forceinline__ f3 icp_scan(const IcpArgs& a, f3 px, int ci, int cj, const float* ly_x, const float* ly_y,
                                       const float* ly_z, const float* ln_x, const float* ln_y, const float* ln_z,
                                       int r0, int c0, int bw)
{
    f3 best{0.0f, 0.0f, 0.0f};
    const int ns = a.n_search;
    float smallest = 1e6f;
    // squared length of the current best: sqrt is monotone, so a candidate
    // whose squared length is not smaller cannot pass `dist < smallest`, and
    // only the others need the sqrt
    float best_sq = __builtin_inff();
    for (int r = ci - ns; r < ci + ns; ++r) {
        for (int c = cj - ns; c < cj + ns; ++c) {
            const unsigned int q = (unsigned int)(r * a.W + c);
            const f3 py = icp_cand<kLds>(ly_x, ly_y, ly_z, r, c, r0, c0, bw, a.y, q);
            if (!kPlane) {
                const f3 dv = sub3(py, px);
                const float sq = dv.x * dv.x + dv.y * dv.y + dv.z * dv.z;
                if (py.z != 0.0f && sq < best_sq) {
                    const float dist = sqrtf(sq);
                    if (dist < smallest && dist < a.max_dist) {
                        smallest = dist;
                        best_sq = sq;
                        best = py;
                    }
                }
            } else {
                const f3 nv = icp_cand<kLds>(ln_x, ln_y, ln_z, r, c, r0, c0, bw, a.yn, q);
                if (py.z != 0.0f && nv.x != 0.0f) {
                    const f3 dv = sub3(py, px);
                    const float d = dv.x * nv.x + dv.y * nv.y + dv.z * nv.z;
                    const float dist = fabsf(d);
                    if (dist < smallest && dist < a.max_dist) {
                        smallest = dist;
                        best = f3{px.x + d * nv.x, px.y + d * nv.y, px.z + d * nv.z};
                    }
                }
            }
        }
    }
    return best;
}

template <bool kPlane>
__global__ __launch_bounds__(kIcpTile * kIcpTile) void k_icp_associate(IcpArgs a)
{
    __shared__ float sy[3][kIcpLdsAlloc];
    __shared__ float sn[kPlane ? 3 : 1][kPlane ? kIcpLdsAlloc : 1];
    __shared__ int box[4];  // row min, row max (exclusive), col min, col max (exclusive)
    const int tiles_x = (a.W + kIcpTile - 1) / kIcpTile;
    const int i = (int)(blockIdx.x / tiles_x) * kIcpTile + (int)(threadIdx.x / kIcpTile);
    const int j = (int)(blockIdx.x % tiles_x) * kIcpTile + (int)(threadIdx.x % kIcpTile);
    const bool in = i < a.H && j < a.W;
    const unsigned int k = in ? (unsigned int)(i * a.W + j) : 0u;
    const int ns = a.n_search;
    bool act = false;
    f3 px{0.0f, 0.0f, 0.0f};
    int ci = 0, cj = 0;
    if (in) {
        const f3 xp = ld3(a.x, k);
        if (xp.z != 0.0f) {
            px = xform(xp, a.T);
            const float cxi = px.y * a.cam.fy / px.z + a.cam.ci;
            const float cxj = px.x * a.cam.fx / px.z + a.cam.cj;
            ci = f2i(cxi);
            cj = f2i(cxj);
            act = ci > ns && ci < a.H - ns && cj > ns && cj < a.W - ns;
        }
    }
    bool stage = false;
    int r0 = 0, c0 = 0, bw = 0;
    if (kIcpLds > 0) {
        if (threadIdx.x == 0) {
            box[0] = 0x7FFFFFFF;
            box[1] = (int)0x80000000u;
            box[2] = 0x7FFFFFFF;
            box[3] = (int)0x80000000u;
        }
        __syncthreads();
        if (act) {
            atomicMin(&box[0], ci - ns);
            atomicMax(&box[1], ci + ns);
            atomicMin(&box[2], cj - ns);
            atomicMax(&box[3], cj + ns);
        }
        __syncthreads();
        // a workgroup without an active pixel leaves the sentinels in place:
        // compare before subtracting (their difference would wrap around).
        // An active pixel has 1 <= ci - ns and ci + ns <= H - 1 (columns
        // alike), so the box lies inside the image.
        const bool any = box[0] < box[1] && box[2] < box[3];
        r0 = any ? box[0] : 0;
        c0 = any ? box[2] : 0;
        const int bh = any ? box[1] - box[0] : 0;
        bw = any ? box[3] - box[2] : 0;
        stage = any && bh * bw <= kIcpLds;
        if (stage) {
            for (int t = threadIdx.x; t < bh * bw; t += kIcpTile * kIcpTile) {
                const int r = r0 + t / bw, c = c0 + t % bw;
                const unsigned int q = (unsigned int)(r * a.W + c);
                const f3 v = ld3(a.y, q);
                sy[0][t] = v.x;
                sy[1][t] = v.y;
                sy[2][t] = v.z;
                if (kPlane) {
                    const f3 nv = ld3(a.yn, q);
                    sn[0][t] = nv.x;
                    sn[kPlane ? 1 : 0][t] = nv.y;
                    sn[kPlane ? 2 : 0][t] = nv.z;
                }
            }
        }
        __syncthreads();
    }
    f3 best{0.0f, 0.0f, 0.0f};
    if (act) {
        if (stage)
            best = icp_scan<true, kPlane>(a, px, ci, cj, sy[0], sy[1], sy[2], sn[0], sn[kPlane ? 1 : 0],
                                          sn[kPlane ? 2 : 0], r0, c0, bw);
        else
            best = icp_scan<false, kPlane>(a, px, ci, cj, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0,
                                           0);
    }
    if (in)
        st3(a.out, k, best);
}

// Sums of one ICP iteration, in double: [0] points of x (z != 0), [1]
// correspondences (x_nearest.z != 0), [2..4] sum x, [5..7] sum x_nearest,
// [8..16] sum x x_nearest^T (row-major), [17] sum |x_nearest - Txy x|^2.
struct MomentArgs {
    const float* x;
    const float* xn;
    double T[12];
    unsigned int n;
    double* partial;  // [kIcpBlocks][kIcpMoments]
};

__device__ __forceinline__ double wave_sum_f64(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(kThreads) void k_icp_moments(MomentArgs a)
{
    double m[kIcpMoments];
#pragma unroll
    for (int q = 0; q < kIcpMoments; ++q)
        m[q] = 0.0;
    for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < a.n; k += gridDim.x * blockDim.x) {
        const f3 x = ld3(a.x, k);
        const f3 y = ld3(a.xn, k);
        if (x.z != 0.0f)
            m[0] += 1.0;
        if (y.z != 0.0f) {
            const double xd[3] = {x.x, x.y, x.z}, yd[3] = {y.x, y.y, y.z};
            m[1] += 1.0;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                m[2 + r] += xd[r];
                m[5 + r] += yd[r];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    m[8 + 3 * r + c] += xd[r] * yd[c];
            }
            double e2 = 0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double t = a.T[4 * r] * xd[0] + a.T[4 * r + 1] * xd[1] + a.T[4 * r + 2] * xd[2] + a.T[4 * r + 3];
                const double e = yd[r] - t;
                e2 += e * e;
            }
            m[17] += e2;
        }
    }
    __shared__ double red[kThreads / 64][kIcpMoments];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kIcpMoments; ++q) {
        const double s = wave_sum_f64(m[q]);
        if (lane == 0)
            red[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < kIcpMoments) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w)
            s += red[w][threadIdx.x];
        a.partial[(size_t)blockIdx.x * kIcpMoments + threadIdx.x] = s;
    }
}

// The blocks' partial sums: one wavefront per moment, each lane summing the
// blocks lane, lane + 64, ... in order, then the fixed wavefront tree.
constexpr int kIcpFinalThreads = 256;

__global__ __launch_bounds__(kIcpFinalThreads) void k_icp_moments_final(const double* partial, int n_blocks,
                                                                        double* out)
{
    const int lane = threadIdx.x & 63;
    for (int q = threadIdx.x >> 6; q < kIcpMoments; q += kIcpFinalThreads / 64) {
        double s = 0.0;
        for (int b = lane; b < n_blocks; b += 64)
            s += partial[(size_t)b * kIcpMoments + q];
        s = wave_sum_f64(s);
        if (lane == 0)
            out[q] = s;
    }
}

// The sums into buf + kIcpBlocks * kIcpMoments (device), with buf's first
// kIcpBlocks * kIcpMoments doubles as the partial sums.
void launch_icp_moments(const float* x, const float* xn, const double* T16, unsigned int n, double* buf, hipStream_t s)
{
    MomentArgs a;
    a.x = x;
    a.xn = xn;
    for (int k = 0; k < 12; ++k)
        a.T[k] = T16[k];
    a.n = n;
    a.partial = buf;
    hipLaunchKernelGGL(k_icp_moments, dim3(kIcpBlocks), dim3(kThreads), 0, s, a);
    hipLaunchKernelGGL(k_icp_moments_final, dim3(1), dim3(kIcpFinalThreads), 0, s, buf, kIcpBlocks,
                       buf + (size_t)kIcpBlocks * kIcpMoments);
}

int icp_grid(size_t n)
{
    const size_t g = (n + kThreads - 1) / kThreads;
    return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

Camera icp_camera(const tsdf_view* v)
{
    // CreatePcdFromDepthView / kokkosICP: ci = resi - cy, cj = cx (float)
    Camera c;
    c.fx = v->fx;
    c.fy = v->fy;
    c.ci = (float)v->height - v->cy;
    c.cj = v->cx;
    return c;
}

bool icp_view_ok(const tsdf_view* v)
{
    // 3 floats a pixel are indexed in 32 bits
    return v && v->height > 0 && v->width > 0 && (int64_t)v->height * v->width < (1ll << 29);
}

bool icp_params_ok(const tsdf_icp_params* p)
{
    return p && (p->type == TSDF_ICP_POINT_TO_POINT || p->type == TSDF_ICP_POINT_TO_PLANE) && p->n_search >= 0 &&
           p->n_search <= 64 && p->n_iterations >= 0;
}

// kokkosICP's workspace (the point and normal maps of both frames and the
// association, the partial sums, a pinned copy of the final sums), kept per
// device until the process ends and grown when a larger image comes: a
// hipFree in every call would wait for the whole device.  Calls on one device
// hold its mutex for their duration.
struct IcpWorkspace {
    std::mutex mu;
    size_t n_cap = 0;  // pixels the point maps hold
    float* buf = nullptr;
    double* mom = nullptr;
    double* mom_host = nullptr;
};
constexpr int kIcpMaxDevices = 64;
IcpWorkspace g_icp_ws[kIcpMaxDevices];

// Makes `w` hold n pixels (the caller holds w.mu); false on allocation failure.
bool icp_workspace(IcpWorkspace& w, size_t n)
{
    if (!w.mom && hipMalloc(&w.mom, (size_t)(kIcpBlocks + 1) * kIcpMoments * sizeof(double)) != hipSuccess) {
        w.mom = nullptr;
        return false;
    }
    if (!w.mom_host && hipHostMalloc(&w.mom_host, kIcpMoments * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        w.mom_host = nullptr;
        return false;
    }
    if (n > w.n_cap) {
        if (w.buf)
            (void)hipFree(w.buf);
        w.buf = nullptr;
        w.n_cap = 0;
        if (hipMalloc(&w.buf, 4 * n * 3 * sizeof(float)) != hipSuccess) {
            w.buf = nullptr;
            return false;
        }
        w.n_cap = n;
    }
    return true;
}

// Device address of `scratch` on each device, looked up once per device (with
// that device current) by the first tsdf_icp_moments_noalloc call there.  The
// mutex also serialises the calls on one device: they share its one array.
struct IcpScratch {
    std::mutex mu;
    double* ptr = nullptr;
};
IcpScratch g_icp_scratch[kIcpMaxDevices];

// 4x4 inverse by cofactors (double, row-major); false when singular.
bool inv4(const double* m, double* out)
{
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (!(det != 0.0) || !std::isfinite(det))
        return false;
    for (int k = 0; k < 16; ++k)
        out[k] = inv[k] / det;
    return true;
}

void mat4_mul(const double* a, const double* b, double* o)
{
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k)
                s += a[4 * r + k] * b[4 * k + c];
            o[4 * r + c] = s;
        }
}

// Orthogonal polar factor Q of a nonsingular 3x3 A (A = Q P, P symmetric
// positive definite) by the scaled Newton iteration X <- (g X + X^-T / g) / 2,
// g = (|X^-1|_F / |X|_F)^(1/2) (Higham, "Computing the polar decomposition",
// 1986).  With A = H^T and H = U S V^T this is V U^T, the rotation of the
// SVD step of the reference (kokkosICP.cpp:173-177).  False when A is
// singular or the iteration does not settle.
bool polar3(const double* A, double* Q)
{
    double X[9];
    for (int k = 0; k < 9; ++k)
        X[k] = A[k];
    for (int it = 0; it < 100; ++it) {
        const double c00 = X[4] * X[8] - X[5] * X[7], c01 = X[5] * X[6] - X[3] * X[8], c02 = X[3] * X[7] - X[4] * X[6];
        const double c10 = X[2] * X[7] - X[1] * X[8], c11 = X[0] * X[8] - X[2] * X[6], c12 = X[1] * X[6] - X[0] * X[7];
        const double c20 = X[1] * X[5] - X[2] * X[4], c21 = X[2] * X[3] - X[0] * X[5], c22 = X[0] * X[4] - X[1] * X[3];
        const double det = X[0] * c00 + X[1] * c01 + X[2] * c02;
        if (!(det != 0.0) || !std::isfinite(det))
            return false;
        // X^-T = cofactor matrix / det
        const double it_[9] = {c00 / det, c01 / det, c02 / det, c10 / det, c11 / det,
                               c12 / det, c20 / det, c21 / det, c22 / det};
        double nx = 0.0, ni = 0.0;
        for (int k = 0; k < 9; ++k) {
            nx += X[k] * X[k];
            ni += it_[k] * it_[k];
        }
        const double g = std::sqrt(std::sqrt(ni / nx));
        double diff = 0.0;
        for (int k = 0; k < 9; ++k) {
            const double v = 0.5 * (g * X[k] + it_[k] / g);
            diff += (v - X[k]) * (v - X[k]);
            X[k] = v;
        }
        if (diff < 1e-30) {
            for (int k = 0; k < 9; ++k)
                Q[k] = X[k];
            return true;
        }
    }
    return false;
}

}  // namespace

extern "C" {

void tsdf_icp_default_params(tsdf_icp_params* p)
{
    if (!p)
        return;
    p->type = TSDF_ICP_POINT_TO_POINT;  // kokkosICP.h:12, postprocessing.cpp:140,198
    p->n_iterations = 10;               // kokkosICP.cpp:49
    p->n_search = 5;                    // :48
    p->max_dist = 0.075f;               // :50
}

tsdf_status tsdf_depth_to_points(const uint16_t* depth, const tsdf_view* cam, float* points, void* stream)
{
    if (!depth || !points || !icp_view_ok(cam))
        return fail(TSDF_BAD_ARG, "NULL buffer or bad view");
    const size_t n = (size_t)cam->height * cam->width;
    hipLaunchKernelGGL(k_depth_to_points, dim3(icp_grid(n)), dim3(kThreads), 0, (hipStream_t)stream, depth,
                       cam->height, cam->width, icp_camera(cam), points);
    HIPC(hipGetLastError());
    return TSDF_OK;
}

tsdf_status tsdf_points_to_normals(const float* points, int32_t height, int32_t width, float* normals, void* stream)
{
    if (!points || !normals || height <= 0 || width <= 0 || (int64_t)height * width >= (1ll << 29))
        return fail(TSDF_BAD_ARG, "NULL buffer or bad size");
    const size_t n = (size_t)height * width;
    hipLaunchKernelGGL(k_normals, dim3(icp_grid(n)), dim3(kThreads), 0, (hipStream_t)stream, points, height, width,
                       normals);
    HIPC(hipGetLastError());
    return TSDF_OK;
}

tsdf_status tsdf_icp_associate(const float* x_points, const float* y_points, const float* y_normals,
                               const float Txy[16], const tsdf_view* cam, const tsdf_icp_params* p, float* x_nearest,
                               void* stream)
{
    if (!x_points || !y_points || !Txy || !x_nearest || !icp_view_ok(cam) || !icp_params_ok(p))
        return fail(TSDF_BAD_ARG, "NULL buffer, bad view or bad parameters");
    if (p->type == TSDF_ICP_POINT_TO_PLANE && !y_normals)
        return fail(TSDF_BAD_ARG, "point-to-plane association needs the normals of y");
    IcpArgs a;
    a.x = x_points;
    a.y = y_points;
    a.yn = y_normals;
    for (int k = 0; k < 12; ++k)
        a.T[k] = Txy[k];
    a.cam = icp_camera(cam);
    a.H = cam->height;
    a.W = cam->width;
    a.n_search = p->n_search;
    a.max_dist = p->max_dist;
    a.type = p->type;
    a.out = x_nearest;
    const int tiles = ((a.W + kIcpTile - 1) / kIcpTile) * ((a.H + kIcpTile - 1) / kIcpTile);
    if (a.type == TSDF_ICP_POINT_TO_PLANE)
        hipLaunchKernelGGL(k_icp_associate<true>, dim3(tiles), dim3(kIcpTile * kIcpTile), 0, (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL(k_icp_associate<false>, dim3(tiles), dim3(kIcpTile * kIcpTile), 0, (hipStream_t)stream, a);
    HIPC(hipGetLastError());
    return TSDF_OK;
}

tsdf_status tsdf_icp_moments(const float* x_points, const float* x_nearest, const double Txy[16], int32_t height,
                             int32_t width, double* moments, void* stream)
{
    if (!x_points || !x_nearest || !Txy || !moments || height <= 0 || width <= 0 ||
        (int64_t)height * width >= (1ll << 29))
        return fail(TSDF_BAD_ARG, "NULL buffer or bad size");
    hipStream_t s = (hipStream_t)stream;
    double* buf = nullptr;
    HIPC(hipMalloc(&buf, (size_t)(kIcpBlocks + 1) * kIcpMoments * sizeof(double)));
    launch_icp_moments(x_points, x_nearest, Txy, (unsigned int)height * (unsigned int)width, buf, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(moments, buf + (size_t)kIcpBlocks * kIcpMoments, kIcpMoments * sizeof(double),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return fail(TSDF_HIP_ERROR, std::string("icp moments: ") + hipGetErrorString(e));
    return TSDF_OK;
}

tsdf_status tsdf_icp_moments_noalloc(const float* x_points, const float* x_nearest, const double Txy[16],
                                     int32_t height, int32_t width, double* moments, void* stream)
{
    if (!x_points || !x_nearest || !Txy || !moments || height <= 0 || width <= 0)
        return fail(TSDF_BAD_ARG, "NULL buffer or non-positive size");
    if ((int64_t)height * width >= (1ll << 29))
        return fail(TSDF_BAD_ARG, "image larger than the no-allocation sums take (at most 2^29 - 1 pixels)");
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kIcpMaxDevices)
        return fail(TSDF_BAD_ARG, "device index beyond the scratch table");
    IcpScratch& sc = g_icp_scratch[dev];
    std::lock_guard<std::mutex> lock(sc.mu);
    if (!sc.ptr) {
        void* p = nullptr;
        const hipError_t ge = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (ge != hipSuccess || !p)
            return fail(TSDF_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(ge));
        sc.ptr = (double*)p;
    }
    hipStream_t s = (hipStream_t)stream;
    // every partial and the final sums are written before they are read, so
    // the previous call's contents need no clearing
    launch_icp_moments(x_points, x_nearest, Txy, (unsigned int)height * (unsigned int)width, sc.ptr, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(moments, sc.ptr + (size_t)kIcpBlocks * kIcpMoments, kIcpMoments * sizeof(double),
                           hipMemcpyDeviceToHost, s);
    // the lock is held until the stream drained: the next call may reuse the array
    const hipError_t se = hipStreamSynchronize(s);
    if (e == hipSuccess)
        e = se;
    if (e != hipSuccess)
        return fail(TSDF_HIP_ERROR, std::string("icp moments (no allocation): ") + hipGetErrorString(e));
    return TSDF_OK;
}

tsdf_status tsdf_icp_release(void)
{
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kIcpMaxDevices)
        return fail(TSDF_BAD_ARG, "device index beyond the ICP workspace table");
    IcpWorkspace& ws = g_icp_ws[dev];
    std::lock_guard<std::mutex> lock(ws.mu);
    // tsdf_icp returns only after its stream drained, so no kernel of a
    // finished call still uses the buffers
    (void)hipFree(ws.buf);
    (void)hipFree(ws.mom);
    if (ws.mom_host)
        (void)hipHostFree(ws.mom_host);
    ws.buf = nullptr;
    ws.mom = nullptr;
    ws.mom_host = nullptr;
    ws.n_cap = 0;
    return TSDF_OK;
}

tsdf_status tsdf_icp(const uint16_t* depth1, const uint16_t* depth2, const double tx[16], const double ty[16],
                     const tsdf_view* cam, const tsdf_icp_params* p, double T_out[16], float* rmse_out,
                     float* fitness_out, void* stream)
{
    if (!depth1 || !depth2 || !tx || !ty || !T_out || !icp_view_ok(cam) || !icp_params_ok(p))
        return fail(TSDF_BAD_ARG, "NULL buffer, bad view or bad parameters");
    double tyi[16], T[16];
    if (!inv4(ty, tyi))
        return fail(TSDF_BAD_ARG, "ty is singular");
    mat4_mul(tyi, tx, T);  // Txy = ty^-1 tx, kokkosICP.cpp:52
    const size_t n = (size_t)cam->height * cam->width;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kIcpMaxDevices)
        return fail(TSDF_BAD_ARG, "device index beyond the ICP workspace table");
    IcpWorkspace& ws = g_icp_ws[dev];
    std::lock_guard<std::mutex> lock(ws.mu);
    // Every exit after the first launch waits for the stream before the lock
    // is released: the kernels read the caller's depth images and write the
    // shared workspace, and the call is synchronous (declared after `lock`,
    // so it runs first on the way out).
    struct StreamDrain {
        hipStream_t s;
        ~StreamDrain() { (void)hipStreamSynchronize(s); }
    } drain{s};
    if (!icp_workspace(ws, n))
        return fail(TSDF_HIP_ERROR, "icp workspace allocation failed");
    float* buf = ws.buf;
    double* mom = ws.mom;
    double* mom_host = ws.mom_host;
    float *x = buf, *y = buf + 3 * n, *yn = buf + 6 * n, *xn = buf + 9 * n;
    tsdf_status r = tsdf_depth_to_points(depth1, cam, x, stream);
    if (r == TSDF_OK)
        r = tsdf_depth_to_points(depth2, cam, y, stream);
    if (r == TSDF_OK)
        r = tsdf_points_to_normals(y, cam->height, cam->width, yn, stream);
    float lowest = 1e6f, best_fit = 0.0f;
    double best[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    for (int it = 0; it < p->n_iterations && r == TSDF_OK; ++it) {
        float Tf[16];
        for (int k = 0; k < 16; ++k)
            Tf[k] = (float)T[k];
        r = tsdf_icp_associate(x, y, yn, Tf, cam, p, xn, stream);
        if (r == TSDF_OK) {
            launch_icp_moments(x, xn, T, (unsigned int)n, mom, s);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(mom_host, mom + (size_t)kIcpBlocks * kIcpMoments, kIcpMoments * sizeof(double),
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                r = fail(TSDF_HIP_ERROR, std::string("icp moments: ") + hipGetErrorString(e));
        }
        const double* m = mom_host;
        if (r != TSDF_OK || m[1] == 0.0)
            break;  // no correspondence: the reference's sums turn to NaN and nothing improves
        const double nc = m[1];
        const double cx[3] = {m[2] / nc, m[3] / nc, m[4] / nc}, cy[3] = {m[5] / nc, m[6] / nc, m[7] / nc};
        const float rmse = std::sqrt((float)(m[17] / nc));
        const float fit = (float)(nc / m[0]);
        if (fit > best_fit)
            best_fit = fit;
        // H = sum (x - cx)(y - cy)^T = sum x y^T - n cx cy^T; the rotation is
        // the polar factor of H^T
        double Ht[9];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                Ht[3 * b + a] = m[8 + 3 * a + b] - nc * cx[a] * cy[b];
        double R[9];
        if (!polar3(Ht, R))
            break;
        for (int a = 0; a < 3; ++a) {
            for (int b = 0; b < 3; ++b)
                T[4 * a + b] = R[3 * a + b];
            T[4 * a + 3] = cy[a] - (R[3 * a] * cx[0] + R[3 * a + 1] * cx[1] + R[3 * a + 2] * cx[2]);
        }
        T[12] = T[13] = T[14] = 0.0;
        T[15] = 1.0;
        if (rmse < lowest) {
            lowest = rmse;
            for (int k = 0; k < 16; ++k)
                best[k] = T[k];
        }
    }
    if (r != TSDF_OK)
        return r;
    // the launches of an n_iterations == 0 call end here too
    const hipError_t se = hipStreamSynchronize(s);
    if (se != hipSuccess)
        return fail(TSDF_HIP_ERROR, std::string("icp: ") + hipGetErrorString(se));
    for (int k = 0; k < 16; ++k)
        T_out[k] = best[k];
    if (rmse_out)
        *rmse_out = lowest;
    if (fitness_out)
        *fitness_out = best_fit;
    return TSDF_OK;
}

}  // extern "C"
