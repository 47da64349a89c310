// su2_draw.hpp -- the SU(2) draw of the device heatbath (heatbath.hip) and its
// random stream, compiled for the host too so that the draw can be checked
// against its exact distribution on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace qcb {

__host__ __device__ inline uint64_t hb_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct hb_rng {
    uint64_t key;
    uint32_t ctr;
    __host__ __device__ inline double unif() {  // (0, 1]
        const uint64_t h = hb_mix(key + 0x9e3779b97f4a7c15ULL * (uint64_t)(++ctr));
        return ((h >> 11) + 1) * 0x1.0p-53;
    }
};

// Draw a in SU(2) (as a quaternion) with density ~ exp(alpha a0) on S^3:
// a0 with density ~ sqrt(1 - a0^2) exp(alpha a0), then a uniform direction.
// For alpha >= creutz_below (kCreutzBelow unless a test forces one method),
// Kennedy-Pendleton (heatbath.cpp:8-36): lambda^2 from two exponentials and a
// cosine, accepted when delta^2 <= 1 - lambda^2.
// Its acceptance falls to zero as alpha -> 0 (0.28 at alpha = 0.5, 1e-3 at
// 0.01), and a hot-started lattice does meet such links, so below the
// threshold a0 comes from Creutz's method instead (M. Creutz, Phys. Rev. D 21
// (1980) 2308): a0 from exp(alpha a0) on [-1, 1] by inverting its CDF,
// accepted with probability sqrt(1 - a0^2).  Both target the same density;
// the acceptance of the pair is above 0.68 for every alpha > 0, so kMaxTries
// bounds the loop without ever being the normal exit.
constexpr int kMaxTries = 64;
constexpr double kCreutzBelow = 2.0;
__host__ __device__ inline bool su2_draw(double alpha, hb_rng& r, double a[4],
                                         double creutz_below = kCreutzBelow) {
    double a0 = 0.0;
    bool ok = false;
    if (alpha >= creutz_below) {
        for (int k = 0; k < kMaxTries; k++) {
            const double delta = r.unif(), r1 = r.unif(), r2 = r.unif(), r3 = r.unif();
            const double cs = cos(2.0 * M_PI * r2);
            const double lsq = -0.5 / alpha * (log(r1) + cs * cs * log(r3));
            if (delta * delta <= 1.0 - lsq) {
                a0 = 1.0 - 2.0 * lsq;
                ok = true;
                break;
            }
        }
    } else {
        const double span = expm1(-2.0 * alpha);  // e^{-2 alpha} - 1, in (-1, 0)
        for (int k = 0; k < kMaxTries; k++) {
            const double u = r.unif(), delta = r.unif();
            const double x = fmin(1.0, fmax(-1.0, 1.0 + log1p((1.0 - u) * span) / alpha));
            if (delta * delta <= 1.0 - x * x) {
                a0 = x;
                ok = true;
                break;
            }
        }
    }
    if (!ok) return false;
    const double rs = sqrt(fmax(0.0, 1.0 - a0 * a0));
    const double phi = 2.0 * M_PI * r.unif();
    const double ct = 2.0 * r.unif() - 1.0;
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    a[0] = a0;
    a[1] = rs * st * cos(phi);
    a[2] = rs * st * sin(phi);
    a[3] = rs * ct;
    return true;
}

}  // namespace qcb
