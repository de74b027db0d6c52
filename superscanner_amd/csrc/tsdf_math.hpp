// tsdf_math.hpp -- scalar arithmetic of the TSDF path shared by every kernel.
//
// These functions define the floating-point contract that makes the GPU
// results bit-identical to the reference algorithm: binary32 in source order
// (the translation unit is compiled with -ffp-contract=off and correctly
// rounded fp32 division / sqrt), the reference's double promotions where its
// C++ source promotes, and the reference's (int) truncations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace tsdf {

constexpr int kVPE = 8;    // voxelPerEdge, GPUVoxelGrid.h:6
constexpr int kVPB = 512;  // voxelPerBlock, GPUVoxelGrid.h:8

// Hash slot: high 32 bits = key (int32 reference hash as bits), low 32 bits =
// brick id.  The all-ones word never occurs as an entry because brick ids stay
// below 2^31.
constexpr unsigned long long kSlotEmpty = ~0ull;
constexpr uint32_t kIdPending = 0xFFFFFFFEu;   // key inserted, id being assigned
constexpr uint32_t kIdOverflow = 0xFFFFFFFDu;  // key inserted, pool was full

struct f3 {
    float x, y, z;
};

// (int) of a float with x86-64 cvttss2si semantics: truncation toward zero,
// INT_MIN for NaN / out of range (the GPU convert would saturate instead).
__host__ __device__ __forceinline__ int f2i(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : (int)0x80000000u;
}

// getPrimeHash, GPUVoxelGrid.h:124-128: 73856093x + 19349669y + 83492791z in
// wrapping 32-bit int arithmetic.  The reference widens the int result to
// unsigned long long by sign extension, so the 32-bit value is the whole key.
__host__ __device__ __forceinline__ uint32_t prime_hash(int x, int y, int z)
{
    return (uint32_t)x * 73856093u + (uint32_t)y * 19349669u + (uint32_t)z * 83492791u;
}

// Bucket scatter for the open-addressing table (murmur3 finaliser).
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// getVBIndeces per axis, GPUVoxelGrid.h:162-190.
__host__ __device__ __forceinline__ int vb_index(float a, float vbl)
{
    return (a >= 0) ? f2i(a / vbl) : f2i(a / vbl - 1.0f);
}

// getVinVBIndeces per axis, GPUVoxelGrid.h:192-221 (p - vbl * vbInd with
// float * int -> float, kokkosFunctions.h:737-740).
__host__ __device__ __forceinline__ int vin_index(float p, int vb, float vbl, float vl)
{
    float q = p - vbl * (float)vb;
    return (q >= 0) ? f2i(q / vl) : f2i(q / vl - 1.0f);
}

// Fast path of the block / voxel index: with rd = 1 / d rounded,
// q = a * rd is within |a/d| * 2^-22 of RN(a / d).  When q is farther than
// max(|q|, 1) * 2^-19 from every integer (ok = true), both RN(a / d) and
// RN(RN(a / d) - 1) lie strictly between the same two integers as q, so the
// reference's truncation (positive branch) and its "- 1 then truncate"
// (negative branch) both equal floor(q).  Otherwise (near an integer,
// |q| >= 2^19, NaN / Inf) the caller must run the reference arithmetic.
__host__ __device__ __forceinline__ int div_fast(float a, float rd, bool& ok)
{
    const float q = a * rd;
    const float f = floorf(q);
    const float fr = q - f;
    const float m = fmaxf(fabsf(q), 1.0f) * 0x1p-19f;
    ok = fr > m && fr < 1.0f - m;
    return ok ? (int)f : 0;
}

// Same value as vb_index(a, d) (and, with a = p - vbl * vb, as vin_index).
__host__ __device__ __forceinline__ int div_index(float a, float d, float rd)
{
    bool ok;
    const int f = div_fast(a, rd, ok);
    if (ok)
        return f;
    return (a >= 0) ? f2i(a / d) : f2i(a / d - 1.0f);
}

// RN32(a / d) exactly, from R = RN64(1 / (double)d) (d a positive normal
// float), for every float a:
//  - t = RN64((double)a * R) is within |a / d| 2^-52 (1 + 2^-54) of a / d;
//  - a / d is never the midpoint mu of two consecutive floats (a = mu d would
//    give a the odd part of mu, at least 2^24 + 1), and otherwise
//    a - mu d is a nonzero multiple of 2^min(e_a, e_mu + e_d) (exponents of
//    the integer significands), which puts a / d at least 2^-50 |a / d| away
//    from every midpoint;
// so no midpoint lies between t and a / d, and RN32(t) = RN32(a / d).
// Quotients below 2^-126 may round differently; every index computed from
// them (vb_index_q, vin_index_q) truncates them to the same integer.
// Infinities and NaN propagate as in the division.
__host__ __device__ __forceinline__ float quot_rn(float a, double R)
{
    return (float)((double)a * R);
}

// getVBIndeces per axis (GPUVoxelGrid.h:162-190) from the exact quotient:
// equal to vb_index(a, d) for every a (R = RN64(1 / d)).
__host__ __device__ __forceinline__ int vb_index_q(float a, double R)
{
    const float q = quot_rn(a, R);
    return (a >= 0) ? f2i(q) : f2i(q - 1.0f);
}

// getVinVBIndeces per axis (GPUVoxelGrid.h:192-221) from the exact quotient:
// equal to vin_index(p, vb, vbl, vl) for every p (Rvl = RN64(1 / vl)).
__host__ __device__ __forceinline__ int vin_index_q(float p, int vb, float vbl, double Rvl)
{
    const float q = p - vbl * (float)vb;
    const float t = quot_rn(q, Rvl);
    return (q >= 0) ? f2i(t) : f2i(t - 1.0f);
}

// Certifies an approximate reciprocal r of z: e = RN(r z - 1) (one fma)
// with |e| <= 2^-22 bounds the relative error of r by 2^-22 (1 + 2^-23),
// whatever produced r (0, Inf, NaN and flushed denormals all fail).
__host__ __device__ __forceinline__ bool rcp_certified(float r, float z)
{
    return fabsf(__builtin_fmaf(r, z, -1.0f)) <= 0x1p-22f;
}

// Pixel index along one image axis as getCoordinate (GPUVoxelGrid.h:234-237)
// and the (int) cast of GPUVoxelGrid.cpp:348-349 give it:
// idx = f2i(RN(RN(n / z) + c0)) with n = RN(c * f).
//
// Fast path with an approximate reciprocal r of z whose relative error the
// caller has certified to be at most 2^-22 (rcp_certified):
//   qa = RN(n * r) is within 0.75 * 2^-21 |qa| of q = RN(n / z), and
//   ra = RN(qa + c0) is within B = 0.75 * 2^-21 |qa| + 0.5 * 2^-22 |ra| of the
//   reference's row = RN(q + c0).
// When ra is farther than m = 2^-21 |qa| + 2^-22 |ra| + 2^-21 from both
// neighbouring integers (m exceeds B by more than the rounding of m, of
// ra - floor(ra) and of 1 - m, and by more than any underflow error), row
// and ra lie in the same open interval (k, k+1), on which truncation is
// constant, so f2i(row) == (int)ra.  The test fails for |ra| >= 2^23, Inf and
// NaN, and the caller then evaluates the reference expression.
__host__ __device__ __forceinline__ bool pix_axis_fast(float n, float r, float c0, int& idx)
{
    const float qa = n * r;
    const float ra = qa + c0;
    const float fr = ra - floorf(ra);
    const float m = __builtin_fmaf(fabsf(qa), 0x1p-21f, __builtin_fmaf(fabsf(ra), 0x1p-22f, 0x1p-21f));
    idx = (int)ra;
    return (fr > m) & (fr < 1.0f - m);
}

// The reference expression of the same index: f2i(RN(RN(n / z) + c0)).
__host__ __device__ __forceinline__ int pix_axis_ref(float n, float z, float c0)
{
    return f2i(n / z + c0);
}

// getVoxelMidPoint per axis, GPUVoxelGrid.h:275-278: the int * float product
// is float, the (v + 0.5) * vl term and the sum are double, rounded to float.
__host__ __device__ __forceinline__ float voxel_mid(int v, int vb, float vbl, float vl)
{
    return (float)((double)((float)vb * vbl) + ((double)v + 0.5) * (double)vl);
}

// Kvec3f scalar product, kokkosFunctions.h:212-214.
__host__ __device__ __forceinline__ float dot3(f3 a, f3 b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// transformPoint, GPUVoxelGrid.h:240-247, with the 3x4 row-major block.
__host__ __device__ __forceinline__ f3 xform(f3 p, const float* E)
{
    f3 o;
    o.x = p.x * E[0] + p.y * E[1] + p.z * E[2] + E[3];
    o.y = p.x * E[4] + p.y * E[5] + p.z * E[6] + E[7];
    o.z = p.x * E[8] + p.y * E[9] + p.z * E[10] + E[11];
    return o;
}

// Camera model of the reference (getPoint GPUVoxelGrid.h:229-232 and
// setIntrinsic GPUVoxelGrid.cpp:253-266): ci = H - cy, cj = cx.
struct Camera {
    float fx, fy, ci, cj;
};

__host__ __device__ __forceinline__ f3 get_point(const Camera& c, int i, int j, float d)
{
    f3 p;
    p.x = ((float)j - c.cj) / c.fx * d;
    p.y = ((float)i - c.ci) / c.fy * d;
    p.z = d;
    return p;
}

// Pose data of one call: E (camera->world) and E^-1 as 3x4 row-major blocks,
// and Ez, the normalised third column of E.
struct Pose {
    float E[12];
    float Ei[12];
    float Ez[3];
    float pad;
};

// Pose of a render view: camera -> world rotation (row-major) and the camera
// centre in world coordinates.
struct ViewPose {
    float R[9];
    float t[3];
};

}  // namespace tsdf
