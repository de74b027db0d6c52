// Host check that the fast pixel index of the integrate kernel
// (tsdf::rcp_certified + tsdf::pix_axis_fast) returns exactly the reference
// index f2i(RN(RN(n / z) + c0)) whenever it claims success.  The reciprocal
// is RN(1/z) perturbed by a random relative error up to +-2^-21 (twice the
// certified bound, so the certificate itself is exercised), plus 0, Inf and
// NaN.  Rows are random and adversarial: the target n / z + c0 is put at and
// one or two ulps around integers.  Exits non-zero on the first mismatch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../superscanner_amd/csrc/tsdf_math.hpp"

int main(int argc, char** argv)
{
    const long n_rand = argc > 1 ? atol(argv[1]) : 10000000L;
    std::mt19937_64 rng(777);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long checked = 0, accepted = 0;
    auto check = [&](float n, float z, float c0, float r) {
        int idx = 0;
        const bool ok = tsdf::rcp_certified(r, z) & tsdf::pix_axis_fast(n, r, c0, idx);
        ++checked;
        if (!ok)
            return;
        ++accepted;
        const int e = tsdf::pix_axis_ref(n, z, c0);
        if (e != idx) {
            std::printf("MISMATCH n=%.9g z=%.9g c0=%.9g r=%.9g ref=%d fast=%d\n", n, z, c0, r, e, idx);
            std::exit(1);
        }
    };
    auto perturbed = [&](float z) {
        const double d = (2.0 * U(rng) - 1.0) * 0x1p-21;
        return (float)((1.0 / (double)z) * (1.0 + d));
    };
    const float c0s[] = {240.5f, 424.5f, 319.5f, 0.0f, 0.5f, 360.5f, 639.5f, -3.25f};
    const float fs[] = {525.0f, 430.0f, 640.0f, 262.5f, 1000.0f};
    for (long k = 0; k < n_rand; ++k) {
        const float c0 = c0s[k % 8], f = fs[k % 5];
        const float z = (float)(0.01 + 12.0 * U(rng)) * (k % 17 == 0 ? -1.0f : 1.0f);
        const float c = (float)((2.0 * U(rng) - 1.0) * 1.5 * std::fabs(z));
        const float n = c * f;
        check(n, z, c0, perturbed(z));
        check(n, z, c0, (float)(1.0 / (double)z));
    }
    // adversarial: n / z + c0 at integers and one / two ulps around them
    for (int row = -3; row <= 1300; ++row) {
        for (int t = 0; t < 40; ++t) {
            const float c0 = c0s[t % 8];
            const float z = (float)(0.05 + 8.0 * U(rng));
            const float q = (float)row - c0;
            float n = (float)((double)q * (double)z);
            for (int s = 0; s < 5; ++s) {
                check(n, z, c0, perturbed(z));
                check(n, z, c0, (float)(1.0 / (double)z));
                check(std::nextafter(n, INFINITY), z, c0, perturbed(z));
                check(std::nextafter(n, -INFINITY), z, c0, perturbed(z));
                n = std::nextafter(n, (t & 1) ? INFINITY : -INFINITY);
            }
        }
    }
    // certificate: non-finite, zero, denormal and grossly wrong reciprocals
    for (float z : {0.0f, -0.0f, 1e-40f, 1e-30f, 3.0f, -2.5f, 1e30f, INFINITY, NAN}) {
        for (float r : {0.0f, INFINITY, -INFINITY, NAN, 1e-30f, 1.0f / z * 1.001f}) {
            check(100.0f, z, 240.5f, r);
            check(-7.0f, z, 0.5f, r);
        }
    }
    std::printf("pix_axis_fast: %ld inputs, %ld accepted by the fast path, all equal to the reference\n",
                checked, accepted);
    return 0;
}
