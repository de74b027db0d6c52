// tsdf_render.hpp -- differentiable trilinear render of the TSDF, its adjoint
// and the Gauss-Newton normal equations of depth tracking (fp32, gfx950).
// Included once, at the end of tsdf_hip.hip.
//
// No reference counterpart: the reference raycast (GPUVoxelGrid.cpp:561-715)
// samples voxel centres and is not differentiable.  The field, march, root
// and adjoint follow the double-precision definition of the oracle
// (render_oracle.c); in brief:
//   field   phi(p) = trilinear interpolation of the sdf at the 8 voxel centres
//           (k + 0.5) vl around p; undefined unless all 8 voxels exist with
//           weight > 0;
//   ray     pixel (i, j) of an h x w view: r = ((j - cx) / fx,
//           (i - (h - cy)) / fy, 1), world point o + lam R r (lam = camera
//           depth; the reference camera convention, GPUVoxelGrid.cpp:253-266);
//   march   from lam = rmin by max(0.5 |phi|, 0.5 vl) / |R r| where phi is
//           defined, 0.5 vl / |R r| where it is not but the block exists, and
//           to the block's exit where the block is not in the map; the first
//           consecutive defined pair with phi > 0 >= phi brackets the surface;
//   root    safeguarded Newton inside the bracket; a hit needs
//           dphi/dlam < -0.05 |grad phi| |R r|;
//   adjoint with s = grad phi . R r at the root p*, for the left
//           perturbation T' = exp(xi^) T, xi = (v, w):
//             dD/dv = -grad phi / s, dD/dw = -(p* x grad phi) / s,
//           and dD/dV_c = -a_c / s for the 8 voxels (trilinear weights a_c).
//
// Every sum over pixels is taken in a fixed order, so that a call repeated on
// the same inputs gives the same bits: a workgroup sums its pixels (a fixed
// shuffle tree, then its wavefronts in order) into its own partial; one
// workgroup per (view, value) sums the partials in fp64, thread t taking
// workgroups t, t + 256, ..., then the same tree.  grad_sdf is a scatter: each
// pixel writes its 8 (voxel index, dL/dV) records at fixed positions, a
// stable radix sort by voxel index keeps the pixel order within a voxel, and
// one thread per voxel adds its run in that order.
#pragma once

namespace {

constexpr int kMaxViews = 256;
constexpr int kRenderMaxSteps = 8192;
constexpr float kRStepFrac = 0.5f;
constexpr float kRMinStepVox = 0.5f;
constexpr float kRGraze = 0.05f;
constexpr int kRootIters = 16;
constexpr int kNormalEq = 29;  // per view: JtJ upper triangle (21), Jtr (6), sum r^2, count
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;  // grad_sdf record of no voxel (sorts last)
// voxel indices brick * 512 + v stay below kNoRecord for brick ids below this
constexpr uint32_t kMaxIndexedBricks = (1u << 23) - 1u;

struct RenderCam {
    int h, w;
    float fx, fy, ci, cj;  // ci = h - cy, cj = cx
    float rmin, dmax;
};

// Brick lookups of one ray: the last block found, reused while the samples
// stay inside it.
struct BrickCache {
    int bx, by, bz, id;
};

__device__ __forceinline__ int cached_brick(const GridView& g, BrickCache& c, int bx, int by, int bz)
{
    if (bx == c.bx && by == c.by && bz == c.bz)
        return c.id;
    const int id = find_brick(g, bx, by, bz);
    c.bx = bx;
    c.by = by;
    c.bz = bz;
    c.id = id;
    return id;
}

struct Sample {
    float phi;
    float grad[3];
    uint32_t idx[8];  // brick * 512 + v of the 8 voxels (corner c = (c&1, c>>1&1, c>>2)),
                      // meaningful for brick ids below kMaxIndexedBricks
    float wt[8];  // trilinear weights
};

// phi (and its gradient) at p; false where the field is undefined.
template <typename T, bool kFull>
__device__ __forceinline__ bool field_sample(const GridView& g, const float* sdf_ovr, float px, float py,
                                             float pz, BrickCache& bc, Sample& smp)
{
    const float rvl = 1.0f / g.vl;
    const float ux = px * rvl - 0.5f, uy = py * rvl - 0.5f, uz = pz * rvl - 0.5f;
    const float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
    if (!(fabsf(fx0) < 1.6e7f && fabsf(fy0) < 1.6e7f && fabsf(fz0) < 1.6e7f))
        return false;  // outside any addressable block (and NaN)
    const int kx = (int)fx0, ky = (int)fy0, kz = (int)fz0;
    const float fr[3] = {ux - fx0, uy - fy0, uz - fz0};
    const int bx[2] = {kx >> 3, (kx + 1) >> 3}, by[2] = {ky >> 3, (ky + 1) >> 3}, bz[2] = {kz >> 3, (kz + 1) >> 3};
    int bid[8];
    float V[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ax = c & 1, ay = (c >> 1) & 1, az = (c >> 2) & 1;
        int id;
        if (ax && bx[1] == bx[0])
            id = bid[c ^ 1];
        else if (ay && by[1] == by[0])
            id = bid[c ^ 2];
        else if (az && bz[1] == bz[0])
            id = bid[c ^ 4];
        else
            id = cached_brick(g, bc, bx[ax], by[ay], bz[az]);
        bid[c] = id;
        if (id < 0)
            return false;
    }
    const T* pay = (const T*)g.payload;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ax = c & 1, ay = (c >> 1) & 1, az = (c >> 2) & 1;
        const int v = ((kx + ax) & 7) + 8 * ((ky + ay) & 7) + 64 * ((kz + az) & 7);
        const uint32_t id = (uint32_t)bid[c] * (uint32_t)kVPB + (uint32_t)v;
        float s, w;
        ld_pair(pay + (size_t)bid[c] * (2 * kVPB), v, s, w);
        if (!(w > 0.0f))
            return false;
        V[c] = sdf_ovr ? sdf_ovr[id] : s;
        smp.idx[c] = id;
    }
    float s = 0.f, g0 = 0.f, g1 = 0.f, g2 = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float ax = (c & 1) ? fr[0] : 1.0f - fr[0];
        const float ay = ((c >> 1) & 1) ? fr[1] : 1.0f - fr[1];
        const float az = ((c >> 2) & 1) ? fr[2] : 1.0f - fr[2];
        s += ax * ay * az * V[c];
        if (kFull) {
            const float sx = (c & 1) ? 1.0f : -1.0f, sy = ((c >> 1) & 1) ? 1.0f : -1.0f,
                        sz = ((c >> 2) & 1) ? 1.0f : -1.0f;
            g0 += sx * ay * az * V[c];
            g1 += ax * sy * az * V[c];
            g2 += ax * ay * sz * V[c];
            smp.wt[c] = ax * ay * az;
        }
    }
    smp.phi = s;
    if (kFull) {
        smp.grad[0] = g0 * rvl;
        smp.grad[1] = g1 * rvl;
        smp.grad[2] = g2 * rvl;
    }
    return true;
}

struct Ray {
    float o[3], d[3], dn;
};

__device__ __forceinline__ Ray make_ray(const ViewPose& P, const RenderCam& cam, int i, int j)
{
    const float rc[3] = {((float)j - cam.cj) / cam.fx, ((float)i - cam.ci) / cam.fy, 1.0f};
    Ray r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.o[a] = P.t[a];
        r.d[a] = P.R[3 * a] * rc[0] + P.R[3 * a + 1] * rc[1] + P.R[3 * a + 2] * rc[2];
    }
    r.dn = sqrtf(r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2]);
    return r;
}

// Depth of the first surface along the ray (0 = none); on a hit, smp holds
// the full sample at the root.
template <typename T>
__device__ float trace(const GridView& g, const float* sdf_ovr, const RenderCam& cam, const Ray& r, Sample& smp)
{
    const float vl = g.vl, vbl = g.vbl;
    BrickCache bc = {0x7FFFFFFF, 0, 0, -1};
    float lam = cam.rmin, prev_lam = 0.f, prev_phi = 0.f;
    bool prev_valid = false;
    for (int step = 0; step < kRenderMaxSteps && lam < cam.dmax; ++step) {
        const float p[3] = {r.o[0] + lam * r.d[0], r.o[1] + lam * r.d[1], r.o[2] + lam * r.d[2]};
        const bool valid = field_sample<T, false>(g, sdf_ovr, p[0], p[1], p[2], bc, smp);
        const float phi = smp.phi;
        if (valid && prev_valid && prev_phi > 0.f && phi <= 0.f) {
            float a = prev_lam, b = lam, fa = prev_phi, fb = phi;
            float x = (fb == 0.f) ? b : b - fb * (b - a) / (fb - fa);
            for (int it = 0; it < kRootIters; ++it) {
                if (!field_sample<T, true>(g, sdf_ovr, r.o[0] + x * r.d[0], r.o[1] + x * r.d[1],
                                           r.o[2] + x * r.d[2], bc, smp))
                    return 0.f;
                const float fx = smp.phi;
                if (fx == 0.f)
                    break;
                if (fx > 0.f)
                    a = x;
                else
                    b = x;
                const float s = smp.grad[0] * r.d[0] + smp.grad[1] * r.d[1] + smp.grad[2] * r.d[2];
                float nx = (s < 0.f) ? x - fx / s : 0.5f * (a + b);
                if (!(nx > a && nx < b))
                    nx = 0.5f * (a + b);
                const float dx = fabsf(nx - x);
                x = nx;
                if (dx <= 2e-7f * (1.0f + x))
                    break;
            }
            if (!field_sample<T, true>(g, sdf_ovr, r.o[0] + x * r.d[0], r.o[1] + x * r.d[1], r.o[2] + x * r.d[2],
                                       bc, smp))
                return 0.f;
            const float s = smp.grad[0] * r.d[0] + smp.grad[1] * r.d[1] + smp.grad[2] * r.d[2];
            const float gn = sqrtf(smp.grad[0] * smp.grad[0] + smp.grad[1] * smp.grad[1] + smp.grad[2] * smp.grad[2]);
            return (s < -kRGraze * gn * r.dn) ? x : 0.f;
        }
        float st;
        if (valid) {
            const float m = kRStepFrac * fabsf(phi);
            st = fmaxf(m, kRMinStepVox * vl) / r.dn;
        } else {
            const int bx = (int)floorf(p[0] * g.rvbl), by = (int)floorf(p[1] * g.rvbl),
                      bz = (int)floorf(p[2] * g.rvbl);
            if (cached_brick(g, bc, bx, by, bz) >= 0) {
                st = kRMinStepVox * vl / r.dn;
            } else {
                const int bk[3] = {bx, by, bz};
                float te = INFINITY;
#pragma unroll
                for (int a2 = 0; a2 < 3; ++a2) {
                    if (r.d[a2] > 0.f)
                        te = fminf(te, ((float)(bk[a2] + 1) * vbl - r.o[a2]) / r.d[a2]);
                    else if (r.d[a2] < 0.f)
                        te = fminf(te, ((float)bk[a2] * vbl - r.o[a2]) / r.d[a2]);
                }
                const float eps = 1e-3f * vl / r.dn;
                st = fmaxf(te - lam + eps, eps);
            }
        }
        prev_lam = lam;
        prev_phi = phi;
        prev_valid = valid;
        lam += st;
    }
    return 0.f;
}

struct RenderArgs {
    const ViewPose* poses;
    RenderCam cam;
    const float* sdf;         // sdf override [n_bricks][512] or null
    float* depth;             // [n][h][w]
    float* normals;           // [n][h][w][3] or null
    const float* grad_depth;  // backward: dL/dD [n][h][w]
    float* grad_pose;         // backward: [n][6]
    float* grad_sdf;          // backward: [n_bricks][512] or null
    float* part;              // per-workgroup partial sums [n][workgroups][values]
    uint32_t* rec_key;        // backward with grad_sdf: [n][h][w][8] voxel index or kNoRecord
    float* rec_val;           // backward with grad_sdf: [n][h][w][8] dL/dV
    const float* depth_obs;   // tracking: observed depth [n][h][w] (m, 0 = none)
    float max_residual;
    double* normal_eq;        // tracking: [n][kNormalEq]
};

// One pixel per lane, 8 x 8 pixels per wavefront, 16 x 16 per workgroup,
// grid.z = view.
__device__ __forceinline__ bool render_pixel(const RenderCam& cam, int& i, int& j)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    i = blockIdx.y * 16 + (wv >> 1) * 8 + (lane >> 3);
    j = blockIdx.x * 16 + (wv & 1) * 8 + (lane & 7);
    return i < cam.h && j < cam.w;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_render(GridView g, RenderArgs a)
{
    int i, j;
    if (!render_pixel(a.cam, i, j))
        return;
    const int view = blockIdx.z;
    const Ray r = make_ray(a.poses[view], a.cam, i, j);
    Sample smp;
    const float lam = trace<T>(g, a.sdf, a.cam, r, smp);
    const size_t u = ((size_t)view * a.cam.h + i) * a.cam.w + j;
    a.depth[u] = lam;
    if (a.normals) {
        float n[3] = {0.f, 0.f, 0.f};
        if (lam > 0.f) {
            const float gn =
                sqrtf(smp.grad[0] * smp.grad[0] + smp.grad[1] * smp.grad[1] + smp.grad[2] * smp.grad[2]);
            n[0] = smp.grad[0] / gn;
            n[1] = smp.grad[1] / gn;
            n[2] = smp.grad[2] / gn;
        }
        a.normals[3 * u + 0] = n[0];
        a.normals[3 * u + 1] = n[1];
        a.normals[3 * u + 2] = n[2];
    }
}

// dD/dxi of a hit pixel at depth lam (false when the field at the root is
// undefined or the ray is not entering the surface there).
template <typename T>
__device__ __forceinline__ bool pixel_jacobian(const GridView& g, const float* sdf_ovr, const Ray& r, float lam,
                                               Sample& smp, float J[6], float& s)
{
    BrickCache bc = {0x7FFFFFFF, 0, 0, -1};
    const float p[3] = {r.o[0] + lam * r.d[0], r.o[1] + lam * r.d[1], r.o[2] + lam * r.d[2]};
    if (!field_sample<T, true>(g, sdf_ovr, p[0], p[1], p[2], bc, smp))
        return false;
    s = smp.grad[0] * r.d[0] + smp.grad[1] * r.d[1] + smp.grad[2] * r.d[2];
    if (!(s < 0.f))
        return false;
    const float* gr = smp.grad;
    const float c[3] = {p[1] * gr[2] - p[2] * gr[1], p[2] * gr[0] - p[0] * gr[2], p[0] * gr[1] - p[1] * gr[0]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        J[k] = -gr[k] / s;
        J[3 + k] = -c[k] / s;
    }
    return true;
}

// Sums `n` floats over the workgroup (result valid in thread 0).
template <int n>
__device__ __forceinline__ void block_sum(float (&v)[n], float* lds)
{
#pragma unroll
    for (int k = 0; k < n; ++k)
        for (int o = 32; o > 0; o >>= 1)
            v[k] += __shfl_xor(v[k], o);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < n; ++k)
            lds[wv * n + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kThreads / 64; ++w)
#pragma unroll
            for (int k = 0; k < n; ++k)
                v[k] += lds[w * n + k];
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_render_backward(GridView g, RenderArgs a)
{
    __shared__ float lds[(kThreads / 64) * 6];
    int i, j;
    const bool in = render_pixel(a.cam, i, j);
    const int view = blockIdx.z;
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (in) {
        const size_t u = ((size_t)view * a.cam.h + i) * a.cam.w + j;
        const float lam = a.depth[u];
        const float gd = a.grad_depth[u];
        uint32_t key[8];
        float val[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            key[c] = kNoRecord;
            val[c] = 0.f;
        }
        if (lam > 0.f && gd != 0.f) {
            const Ray r = make_ray(a.poses[view], a.cam, i, j);
            Sample smp;
            float J[6], s;
            if (pixel_jacobian<T>(g, a.sdf, r, lam, smp, J, s)) {
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    acc[k] = gd * J[k];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    key[c] = smp.idx[c];
                    val[c] = gd * (-smp.wt[c] / s);
                }
            }
        }
        if (a.rec_key) {
            uint4* kq = (uint4*)(a.rec_key + u * 8);
            float4* vq = (float4*)(a.rec_val + u * 8);
            kq[0] = make_uint4(key[0], key[1], key[2], key[3]);
            kq[1] = make_uint4(key[4], key[5], key[6], key[7]);
            vq[0] = make_float4(val[0], val[1], val[2], val[3]);
            vq[1] = make_float4(val[4], val[5], val[6], val[7]);
        }
    }
    block_sum<6>(acc, lds);
    if (threadIdx.x == 0) {
        const size_t wg = ((size_t)view * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            a.part[wg * 6 + k] = acc[k];
    }
}

// out[view][k] = the sum of the per-workgroup partials part[view][w][k] over
// the nwg workgroups of a view, in fp64, in a fixed order.  grid = (nval,
// n_views), one workgroup each.
template <typename Out>
__global__ __launch_bounds__(kThreads) void k_reduce_parts(const float* part, int nwg, int nval, Out* out)
{
    __shared__ double lds[kThreads / 64];
    const int k = blockIdx.x, view = blockIdx.y;
    double s = 0.0;
    for (int w = threadIdx.x; w < nwg; w += kThreads)
        s += (double)part[((size_t)view * nwg + w) * nval + k];
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        lds[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = lds[0];
        for (int w = 1; w < kThreads / 64; ++w)
            t += lds[w];
        out[(size_t)view * nval + k] = (Out)t;
    }
}

// grad[key] += the values of key's run in the sorted records, summed in fp64
// in record order (pixel order, the sort being stable); one thread per run.
// The records of no voxel sort last, so a thread stops at the first one; a
// run is read four records per round trip.
__global__ __launch_bounds__(kThreads) void k_segment_add(const uint32_t* key, const float* val, size_t n,
                                                          float* grad)
{
    for (size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x; p < n; p += (size_t)gridDim.x * kThreads) {
        const uint32_t k = key[p];
        if (k == kNoRecord)
            break;
        if (p > 0 && key[p - 1] == k)
            continue;
        double t = 0.0;
        for (size_t q = p; q < n; q += 4) {
            uint32_t kk[4];
            float vv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool in = q + u < n;
                kk[u] = in ? key[q + u] : kNoRecord;
                vv[u] = in ? val[q + u] : 0.f;
            }
            bool more = true;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                more = more && kk[u] == k;
                t = more ? t + (double)vv[u] : t;
            }
            if (!more)
                break;
        }
        grad[k] = (float)((double)grad[k] + t);
    }
}

// Gauss-Newton normal equations of depth tracking: for every pixel where
// the render hits and the observation is valid with |D - D_obs| <
// max_residual, J = dD/dxi and r = D - D_obs; per view sum J J^T, J r, r^2
// and the pixel count (accumulated in fp64).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_track(GridView g, RenderArgs a)
{
    __shared__ float lds[(kThreads / 64) * kNormalEq];
    int i, j;
    const bool in = render_pixel(a.cam, i, j);
    const int view = blockIdx.z;
    float acc[kNormalEq];
#pragma unroll
    for (int k = 0; k < kNormalEq; ++k)
        acc[k] = 0.f;
    if (in) {
        const Ray r = make_ray(a.poses[view], a.cam, i, j);
        Sample smp;
        const float lam = trace<T>(g, nullptr, a.cam, r, smp);
        const size_t u = ((size_t)view * a.cam.h + i) * a.cam.w + j;
        if (a.depth)
            a.depth[u] = lam;
        const float dobs = a.depth_obs[u];
        const float res = lam - dobs;
        float J[6], s;
        if (lam > 0.f && dobs > 0.f && fabsf(res) < a.max_residual && pixel_jacobian<T>(g, nullptr, r, lam, smp, J, s)) {
            int k = 0;
#pragma unroll
            for (int p = 0; p < 6; ++p)
#pragma unroll
                for (int q = p; q < 6; ++q)
                    acc[k++] = J[p] * J[q];
#pragma unroll
            for (int p = 0; p < 6; ++p)
                acc[21 + p] = J[p] * res;
            acc[27] = res * res;
            acc[28] = 1.0f;
        }
    }
    block_sum<kNormalEq>(acc, lds);
    if (threadIdx.x == 0) {
        const size_t wg = ((size_t)view * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
#pragma unroll
        for (int k = 0; k < kNormalEq; ++k)
            a.part[wg * kNormalEq + k] = acc[k];
    }
}

__global__ void k_fill_f64(double* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

tsdf_status render_setup(tsdf_grid* g, int32_t n_views, const float* E, const tsdf_view* cam, hipStream_t s,
                         RenderArgs& a, dim3& grid)
{
    if (!g || !E || !cam || n_views < 0)
        return fail(TSDF_BAD_ARG, "NULL argument or negative view count");
    if (n_views > kMaxViews)
        return fail(TSDF_BAD_ARG, "more than 256 views in one call");
    if (cam->height <= 0 || cam->width <= 0 || cam->height > 16384 || cam->width > 16384 ||
        !(cam->fx > 0.f) || !(cam->fy > 0.f) || !std::isfinite(cam->cx) || !std::isfinite(cam->cy))
        return fail(TSDF_BAD_ARG, "bad view size or intrinsics");
    for (int k = 0; k < n_views; ++k)
        if (!pose_ok(E + 16 * (size_t)k))
            return fail(TSDF_BAD_POSE, "pose of a view is not a finite rigid transform");
    if (g->render_staged)
        HIPC(hipEventSynchronize(g->render_done));
    for (int k = 0; k < n_views; ++k) {
        ViewPose& P = g->view_host[k];
        const float* e = E + 16 * (size_t)k;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c)
                P.R[3 * r + c] = e[4 * r + c];
            P.t[r] = e[4 * r + 3];
        }
    }
    if (n_views > 0) {
        HIPC(hipMemcpyAsync(g->view_dev, g->view_host, n_views * sizeof(ViewPose), hipMemcpyHostToDevice, s));
        HIPC(hipEventRecord(g->render_done, s));
        g->render_staged = true;
    }
    memset(&a, 0, sizeof(a));
    a.poses = g->view_dev;
    a.cam.h = cam->height;
    a.cam.w = cam->width;
    a.cam.fx = cam->fx;
    a.cam.fy = cam->fy;
    a.cam.ci = (float)cam->height - cam->cy;
    a.cam.cj = cam->cx;
    a.cam.rmin = g->p.min_depth;
    a.cam.dmax = g->p.max_ray_length;
    grid = dim3((cam->width + 15) / 16, (cam->height + 15) / 16, n_views);
    return TSDF_OK;
}

// Grows the grid's reduction scratch to at least `part` partial sums,
// `rec` grad_sdf records (two copies: the sort's input and output) and `tmp`
// bytes of sort temporary.  Calls that overlap on one grid share it.
tsdf_status ensure_reduce(tsdf_grid* g, size_t part, size_t rec, size_t tmp)
{
    if (g->red_part_cap >= part && g->rec_cap >= rec && g->sort_tmp_cap >= tmp)
        return TSDF_OK;
    HIPC(hipDeviceSynchronize());  // work queued before may still use the old buffers
    if (g->red_part_cap < part) {
        (void)hipFree(g->red_part);
        g->red_part = nullptr;
        g->red_part_cap = 0;
        HIPC(hipMalloc(&g->red_part, part * sizeof(float)));
        g->red_part_cap = part;
    }
    if (g->rec_cap < rec) {
        for (int k = 0; k < 2; ++k) {
            (void)hipFree(g->rec_key[k]);
            (void)hipFree(g->rec_val[k]);
            g->rec_key[k] = nullptr;
            g->rec_val[k] = nullptr;
        }
        g->rec_cap = 0;
        for (int k = 0; k < 2; ++k) {
            HIPC(hipMalloc(&g->rec_key[k], rec * sizeof(uint32_t)));
            HIPC(hipMalloc(&g->rec_val[k], rec * sizeof(float)));
        }
        g->rec_cap = rec;
    }
    if (g->sort_tmp_cap < tmp) {
        (void)hipFree(g->sort_tmp);
        g->sort_tmp = nullptr;
        g->sort_tmp_cap = 0;
        HIPC(hipMalloc(&g->sort_tmp, tmp));
        g->sort_tmp_cap = tmp;
    }
    return TSDF_OK;
}

// exp of the se(3) vector xi = (v, w) as a row-major 4x4 (fp64):
// R = I + A K + B K^2, t = (I + B K + C K^2) v with K = [w]x, th = |w|,
// A = sin th / th, B = (1 - cos th) / th^2, C = (th - sin th) / th^3.
void se3_exp(const double xi[6], double T[16])
{
    const double wx = xi[3], wy = xi[4], wz = xi[5];
    const double th2 = wx * wx + wy * wy + wz * wz, th = std::sqrt(th2);
    double A, B, C;
    if (th < 1e-5) {  // series to O(th^4)
        A = 1.0 - th2 / 6.0;
        B = 0.5 - th2 / 24.0;
        C = 1.0 / 6.0 - th2 / 120.0;
    } else {
        A = std::sin(th) / th;
        B = (1.0 - std::cos(th)) / th2;
        C = (th - std::sin(th)) / (th2 * th);
    }
    const double K[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    double K2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int k = 0; k < 3; ++k)
                s += K[3 * r + k] * K[3 * k + c];
            K2[3 * r + c] = s;
        }
    double V[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double id = r == c ? 1.0 : 0.0;
            T[4 * r + c] = id + A * K[3 * r + c] + B * K2[3 * r + c];
            V[3 * r + c] = id + B * K[3 * r + c] + C * K2[3 * r + c];
        }
    for (int r = 0; r < 3; ++r)
        T[4 * r + 3] = V[3 * r] * xi[0] + V[3 * r + 1] * xi[1] + V[3 * r + 2] * xi[2];
    T[12] = T[13] = T[14] = 0.0;
    T[15] = 1.0;
}

}  // namespace

extern "C" {

tsdf_status tsdf_se3_gauss_newton_step(const double JtJ[21], const double Jtr[6], const double T_in[16],
                                       double T_out[16], double delta[6])
{
    if (!JtJ || !Jtr || !T_in || !T_out)
        return fail(TSDF_BAD_ARG, "NULL argument");
    double H[36];
    int k = 0;
    for (int p = 0; p < 6; ++p)
        for (int q = p; q < 6; ++q) {
            H[6 * p + q] = H[6 * q + p] = JtJ[k++];
        }
    // Cholesky H = L L^T
    double L[36] = {0};
    for (int j = 0; j < 6; ++j) {
        double d = H[6 * j + j];
        for (int m = 0; m < j; ++m)
            d -= L[6 * j + m] * L[6 * j + m];
        if (!(d > 0.0) || !std::isfinite(d))
            return fail(TSDF_BAD_ARG, "normal equations are not positive definite");
        L[6 * j + j] = std::sqrt(d);
        for (int i = j + 1; i < 6; ++i) {
            double s = H[6 * i + j];
            for (int m = 0; m < j; ++m)
                s -= L[6 * i + m] * L[6 * j + m];
            L[6 * i + j] = s / L[6 * j + j];
        }
    }
    double y[6], x[6];
    for (int i = 0; i < 6; ++i) {  // L y = -Jtr
        double s = -Jtr[i];
        for (int m = 0; m < i; ++m)
            s -= L[6 * i + m] * y[m];
        y[i] = s / L[6 * i + i];
    }
    for (int i = 5; i >= 0; --i) {  // L^T x = y
        double s = y[i];
        for (int m = i + 1; m < 6; ++m)
            s -= L[6 * m + i] * x[m];
        x[i] = s / L[6 * i + i];
    }
    double D[16], out[16];
    se3_exp(x, D);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double s = 0;
            for (int m = 0; m < 4; ++m)
                s += D[4 * r + m] * T_in[4 * m + c];
            out[4 * r + c] = s;
        }
    for (int i = 0; i < 16; ++i)
        T_out[i] = out[i];
    if (delta)
        for (int i = 0; i < 6; ++i)
            delta[i] = x[i];
    return TSDF_OK;
}

tsdf_status tsdf_render(tsdf_grid* g, int32_t n_views, const float* E, const tsdf_view* cam, const float* sdf,
                        float* depth, float* normals, void* stream)
{
    if (!depth)
        return fail(TSDF_BAD_ARG, "depth output is NULL");
    DeviceScope ds(g ? g->p.device : 0);
    hipStream_t s = (hipStream_t)stream;
    RenderArgs a;
    dim3 grid;
    tsdf_status st = render_setup(g, n_views, E, cam, s, a, grid);
    if (st != TSDF_OK || n_views == 0)
        return st;
    if (sdf && g->p.n_bricks > kMaxIndexedBricks)
        return fail(TSDF_BAD_ARG, "an sdf override addresses at most 2^23 - 1 bricks");
    a.sdf = sdf;
    a.depth = depth;
    a.normals = normals;
    if (g->p.storage == TSDF_STORAGE_F16)
        hipLaunchKernelGGL(k_render<__half>, grid, dim3(kThreads), 0, s, view(g), a);
    else
        hipLaunchKernelGGL(k_render<float>, grid, dim3(kThreads), 0, s, view(g), a);
    HIPC(hipGetLastError());
    return TSDF_OK;
}

tsdf_status tsdf_render_backward(tsdf_grid* g, int32_t n_views, const float* E, const tsdf_view* cam,
                                 const float* sdf, const float* depth, const float* grad_depth, float* grad_pose,
                                 float* grad_sdf, void* stream)
{
    if (!depth || !grad_depth || !grad_pose)
        return fail(TSDF_BAD_ARG, "depth, grad_depth or grad_pose is NULL");
    DeviceScope ds(g ? g->p.device : 0);
    hipStream_t s = (hipStream_t)stream;
    RenderArgs a;
    dim3 grid;
    tsdf_status st = render_setup(g, n_views, E, cam, s, a, grid);
    if (st != TSDF_OK || n_views == 0)
        return st;
    if ((sdf || grad_sdf) && g->p.n_bricks > kMaxIndexedBricks)
        return fail(TSDF_BAD_ARG, "an sdf override or grad_sdf addresses at most 2^23 - 1 bricks");
    const size_t nwg = (size_t)grid.x * grid.y;
    const size_t n_rec = grad_sdf ? (size_t)n_views * cam->height * cam->width * 8 : 0;
    if (n_rec > (size_t)0x7FFFFFFF)
        return fail(TSDF_BAD_ARG, "grad_sdf over more than 2^31 - 1 records (8 per pixel of all views)");
    // radix bits: kNoRecord's low bits sort after every voxel index
    int bits = 1;
    while (((1ull << bits) - 1ull) < (unsigned long long)g->p.n_bricks * kVPB)
        ++bits;
    size_t tmp = 0;
    if (n_rec)
        HIPC(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, g->rec_key[0], g->rec_key[1], g->rec_val[0],
                                                g->rec_val[1], (int)n_rec, 0, bits, s));
    st = ensure_reduce(g, nwg * n_views * 6, n_rec, tmp);
    if (st != TSDF_OK)
        return st;
    a.sdf = sdf;
    a.depth = (float*)depth;
    a.grad_depth = grad_depth;
    a.grad_pose = grad_pose;
    a.grad_sdf = grad_sdf;
    a.part = g->red_part;
    a.rec_key = n_rec ? g->rec_key[0] : nullptr;
    a.rec_val = n_rec ? g->rec_val[0] : nullptr;
    if (g->p.storage == TSDF_STORAGE_F16)
        hipLaunchKernelGGL(k_render_backward<__half>, grid, dim3(kThreads), 0, s, view(g), a);
    else
        hipLaunchKernelGGL(k_render_backward<float>, grid, dim3(kThreads), 0, s, view(g), a);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_parts<float>, dim3(6, n_views), dim3(kThreads), 0, s, g->red_part, (int)nwg, 6,
                       grad_pose);
    HIPC(hipGetLastError());
    if (n_rec) {
        HIPC(hipcub::DeviceRadixSort::SortPairs(g->sort_tmp, tmp, g->rec_key[0], g->rec_key[1], g->rec_val[0],
                                                g->rec_val[1], (int)n_rec, 0, bits, s));
        const size_t nb = std::min((n_rec + kThreads - 1) / kThreads, (size_t)g->n_cu * 16);
        hipLaunchKernelGGL(k_segment_add, dim3((unsigned)nb), dim3(kThreads), 0, s, g->rec_key[1], g->rec_val[1],
                           n_rec, grad_sdf);
        HIPC(hipGetLastError());
    }
    return TSDF_OK;
}

tsdf_status tsdf_track_normal_eq(tsdf_grid* g, int32_t n_views, const float* E, const tsdf_view* cam,
                                 const float* depth_obs, float max_residual, double* normal_eq, float* depth,
                                 void* stream)
{
    if (!depth_obs || !normal_eq)
        return fail(TSDF_BAD_ARG, "depth_obs or normal_eq is NULL");
    if (!(max_residual > 0.f))
        return fail(TSDF_BAD_ARG, "max_residual must be positive");
    DeviceScope ds(g ? g->p.device : 0);
    hipStream_t s = (hipStream_t)stream;
    RenderArgs a;
    dim3 grid;
    tsdf_status st = render_setup(g, n_views, E, cam, s, a, grid);
    if (st != TSDF_OK || n_views == 0)
        return st;
    const size_t nwg = (size_t)grid.x * grid.y;
    st = ensure_reduce(g, nwg * n_views * kNormalEq, 0, 0);
    if (st != TSDF_OK)
        return st;
    a.depth = depth;
    a.depth_obs = depth_obs;
    a.max_residual = max_residual;
    a.normal_eq = normal_eq;
    a.part = g->red_part;
    if (g->p.storage == TSDF_STORAGE_F16)
        hipLaunchKernelGGL(k_track<__half>, grid, dim3(kThreads), 0, s, view(g), a);
    else
        hipLaunchKernelGGL(k_track<float>, grid, dim3(kThreads), 0, s, view(g), a);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_parts<double>, dim3(kNormalEq, n_views), dim3(kThreads), 0, s, g->red_part,
                       (int)nwg, kNormalEq, normal_eq);
    HIPC(hipGetLastError());
    return TSDF_OK;
}

}  // extern "C"
