// Point-dipole coupling tensor of two spins on the device (the reference's
// dipole(), source/misc/misc.cpp:11-36), shared by the device term builder
// and the factored-propagator kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace oops {

constexpr double kCoupHbar = 1.0545718e-34;
constexpr double kCoupMu0 = 1.25663706143592e-06;
constexpr double kCoupPi = 3.14159265358979323846;
constexpr double kCoupDistEps = 1e-10;   // Angstrom: coincident spins do not couple

// res[3a + b] = pref (delta_ab - 3 n_a n_b) for the unit vector n from spin 2
// to spin 1, with the host's operation order (distance ((dx^2 + dz^2) + dy^2),
// prefactor ((hbar mu0) / (4 pi)) (g1 g2) / (d0 d0 d0), d0 in metres) and no
// contraction into fused multiply-adds, so that it equals the host's value
// bit for bit.
__device__ inline void dipole9(const double* r1, const double* r2, double g1, double g2, double* res) {
#pragma clang fp contract(off)
    const double dx = r1[0] - r2[0], dy = r1[1] - r2[1], dz = r1[2] - r2[2];
    double acc1 = dx * dx;
    const double acc2 = dy * dy;
    acc1 = acc1 + dz * dz;
    const double d = sqrt(acc1 + acc2);
    if (d <= kCoupDistEps) {
        for (int q = 0; q < 9; ++q) res[q] = 0.0;
        return;
    }
    const double n[3] = {(r1[0] - r2[0]) / d, (r1[1] - r2[1]) / d, (r1[2] - r2[2]) / d};
    const double d0 = d * 1e-10;
    const double pref = kCoupHbar * kCoupMu0 / (4.0 * kCoupPi) * (g1 * g2) / (d0 * d0 * d0);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) res[3 * a + b] = pref * ((a == b ? 1.0 : 0.0) - 3.0 * n[a] * n[b]);
}

}  // namespace oops
