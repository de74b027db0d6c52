// Host check that tsdf::div_index (approximate quotient + certified margin)
// and tsdf::vb_index_q / vin_index_q (quotient through one double product)
// return exactly the reference block / voxel index of tsdf::vb_index /
// vin_index, and quot_rn exactly the IEEE quotient, for random and
// adversarial inputs.  Built with hipcc as host code and run by
// tests/test_fast_index.py; exits non-zero on the first mismatch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <random>

#include "../../superscanner_amd/csrc/tsdf_math.hpp"

static float next_up(float x) { return std::nextafter(x, INFINITY); }
static float next_dn(float x) { return std::nextafter(x, -INFINITY); }

int main(int argc, char** argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 20000000L;
    std::mt19937_64 rng(12345);
    const float dens[] = {0.0004f, 0.0016f, 0.002f, 0.004f, 0.005f, 0.01f, 0.016f, 0.02f,
                          0.032f, 0.04f, 0.08f, 0.16f, 0.37f, 1.0f};
    long checked = 0, slow = 0;
    auto check = [&](float a, float d, float rd) {
        const int e = tsdf::vb_index(a, d);
        const int g = tsdf::div_index(a, d, rd);
        const double R = 1.0 / (double)d;
        const int qi = tsdf::vb_index_q(a, R);
        ++checked;
        if (e != g || e != qi) {
            std::printf("MISMATCH a=%.9g (0x%08x) d=%.9g exact=%d fast=%d quot=%d\n", a, *(unsigned*)&a, d, e, g, qi);
            std::exit(1);
        }
        // the quotient itself is the correctly rounded a / d
        const float qa = a / d, qb = tsdf::quot_rn(a, R);
        if (!(qa == qb || (std::isnan(qa) && std::isnan(qb)) || (std::fabs(qa) < 0x1p-126f && std::fabs(qb) < 0x1p-126f))) {
            std::printf("QUOTIENT a=%.9g d=%.9g a/d=%.9g quot=%.9g\n", a, d, qa, qb);
            std::exit(1);
        }
        // the voxel index within a block, for the block of a and its neighbours
        if (std::isfinite(a) && std::fabs(a) < 1e6f) {
            const float vl = d / 8.0f;
            const double Rvl = 1.0 / (double)vl;
            for (int db = -1; db <= 1; ++db) {
                const int vb = e + db;
                const int ve = tsdf::vin_index(a, vb, d, vl), vq = tsdf::vin_index_q(a, vb, d, Rvl);
                if (ve != vq) {
                    std::printf("VIN MISMATCH a=%.9g vb=%d vl=%.9g exact=%d quot=%d\n", a, vb, vl, ve, vq);
                    std::exit(1);
                }
            }
        }
        const float q = a * rd, f = std::floor(q), fr = q - f, m = std::fmax(std::fabs(q), 1.f) * 0x1p-19f;
        slow += !(fr > m && fr < 1.f - m);
    };
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (float d0 : dens) {
        for (int mul : {1, 8}) {
            const float d = (float)mul * d0;
            const float rd = 1.0f / d;
            // random magnitudes up to 100 m
            for (long k = 0; k < n / 28; ++k) {
                const double s = std::pow(10.0, 3.0 * std::fabs(U(rng)) - 1.0);
                check((float)(U(rng) * s), d, rd);
            }
            // adversarial: at and around every integer multiple of d, both signs
            for (int i = -3000; i <= 3000; ++i) {
                float a = (float)i * d;
                for (int s = 0; s < 6; ++s) {
                    check(a, d, rd);
                    check(next_up(a), d, rd);
                    check(next_dn(a), d, rd);
                    a = next_up(a);
                }
                const float b = (float)((double)i * (double)d);
                check(b, d, rd);
                check(next_up(b), d, rd);
                check(next_dn(b), d, rd);
            }
        }
    }
    for (float a : {0.0f, -0.0f, 1e-30f, -1e-30f, 1e-45f, -1e-45f, INFINITY, -INFINITY, NAN, 3e9f, -3e9f})
        check(a, 0.04f, 1.0f / 0.04f);
    // every float of both signs in [2^-20, 2^8) m: quot_rn equals the IEEE
    // quotient for the voxel and block lengths of the benchmark workloads
    if (argc > 2 && std::strcmp(argv[2], "sweep") == 0) {
        long swept = 0;
        for (float d : {0.002f, 0.004f, 0.005f, 0.016f, 0.032f, 0.04f}) {
            const double R = 1.0 / (double)d;
            const float lo = 0x1p-20f, hi = 0x1p8f;
            uint32_t b0, b1;
            std::memcpy(&b0, &lo, 4);
            std::memcpy(&b1, &hi, 4);
            for (uint32_t b = b0; b < b1; ++b)
                for (uint32_t sgn : {0u, 0x80000000u}) {
                    const uint32_t bb = b | sgn;
                    float a;
                    std::memcpy(&a, &bb, 4);
                    if (tsdf::quot_rn(a, R) != a / d) {
                        std::printf("SWEEP MISMATCH a=%.9g d=%.9g\n", a, d);
                        std::exit(1);
                    }
                    ++swept;
                }
        }
        std::printf("sweep: %ld quotients equal to the IEEE division\n", swept);
    }
    // every uint16 depth: quot_rn(u, RN64(1 / 1000)) is the reference's
    // (float)u / 1000.0f (getPoint's metres, the depth-gather update)
    for (uint32_t u = 0; u <= 0xFFFFu; ++u) {
        const float a = (float)u;
        if (tsdf::quot_rn(a, 1.0 / 1000.0) != a / 1000.0f) {
            std::printf("DEPTH MISMATCH u=%u\n", u);
            std::exit(1);
        }
    }
    std::printf("depth: 65536 metre values equal to the IEEE division\n");
    std::printf("div_index: %ld inputs agree with the reference arithmetic (%ld took the exact path)\n",
                checked, slow);
    return 0;
}
