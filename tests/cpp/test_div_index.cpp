// Host check that tsdf::div_index (approximate quotient + certified margin)
// returns exactly the reference block / voxel index of tsdf::vb_index for
// random and adversarial inputs.  Built with hipcc as host code and run by
// tests/test_fast_index.py; exits non-zero on the first mismatch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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
        ++checked;
        if (e != g) {
            std::printf("MISMATCH a=%.9g (0x%08x) d=%.9g exact=%d fast=%d\n", a, *(unsigned*)&a, d, e, g);
            std::exit(1);
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
    std::printf("div_index: %ld inputs agree with the reference arithmetic (%ld took the exact path)\n",
                checked, slow);
    return 0;
}
