// gf_inflow.h -- hydrograph inflow boundary state (device).
//
// Restates the discharge form of hydrograph/hydrograh_module.f90 as used by
// the mthbc = 0 branch of fc2d_geoclaw_bc2 (fortran_source/
// geoflood_bc2_fort.f90:52-102 and the right/bottom/top twins): the
// interpolated discharge per unit width hu0 and, for a wet boundary cell, the
// ghost depth that satisfies the Riemann invariant (Newton, Fr = 0.5 start)
// or, when that depth exceeds the cell's, the two-shock relation (Fr = 1
// start).  Same operation order as the reference (compiled without FMA
// contraction), so it agrees with the CPU oracle to rounding of pow/sqrt.
#pragma once

#include <hip/hip_runtime.h>

#include "gf_pow.h"

namespace gf {

// hydrograh_module.f90:186-196: clamp outside the table, else the first
// interval [time(i), time(i+1)] holding t.  Evaluated by a whole wavefront
// (every lane must call it, with the same arguments; every lane returns the
// value): lane l tests interval base + l for 64 intervals at a time, and the
// lowest lane that holds t is the serial search's first interval, so the
// value is the serial formula's on the same interval.
__device__ inline double hydro_interpolate(const double* __restrict__ tt, const double* __restrict__ hu,
                                           int n, double t)
{
    if (t < tt[0]) return hu[0];
    if (t > tt[n - 1]) return hu[n - 1];
    const int lane = (int)(threadIdx.x & 63u);
    for (int base = 0; base < n - 1; base += 64) {
        const int i = base + lane;
        const bool hit = i < n - 1 && t >= tt[i] && t <= tt[i + 1];
        const unsigned long long m = __ballot(hit);
        if (m) {
            const int f = base + __builtin_ctzll(m);
            return hu[f] + (hu[f + 1] - hu[f]) / (tt[f + 1] - tt[f]) * (t - tt[f]);
        }
    }
    return 0.0;
}

// The reference's tol = 1.0e-6 is a default-real literal.
constexpr double HYDRO_TOL = (double)1.0e-6f;
constexpr int HYDRO_MAXITER = 100;

// Riemann_invariants, discharge branch (hydrograh_module.f90:213-242).
__device__ inline double riemann_invariant_depth(double hu0, double h1, double hu1, double g)
{
    const double u1 = hu1 / h1;
    const double Fr = 0.5;
    double h0 = pow_2_3(hu0 / sqrt(g) * Fr);
    for (int it = 0; it < HYDRO_MAXITER; it++) {
        const double fxn = hu0 / h0 - 2.0 * sqrt(g * h0) - u1 + 2.0 * sqrt(g * h1);
        if (fabs(fxn) < HYDRO_TOL) break;
        const double dfxn = -(hu0 / (h0 * h0)) - sqrt(g / h0);
        h0 = h0 - fxn / dfxn;
    }
    return h0;
}

// two_shock, discharge branch (hydrograh_module.f90:274-303).
__device__ inline double two_shock_depth(double hu0, double h1, double hu1, double g)
{
    const double u1 = hu1 / h1;
    const double Fr = 1.0;
    double h0 = pow_2_3(hu0 / sqrt(g) * Fr);
    for (int it = 0; it < HYDRO_MAXITER; it++) {
        const double fxn = hu0 / h0 - u1 - (h0 - h1) * sqrt((g / 2.0) * (1.0 / h0 + 1.0 / h1));
        if (fabs(fxn) < HYDRO_TOL) break;
        const double num = sqrt(g * (1.0 / h0 + 1.0 / h1));
        const double deno = 2.0 * sqrt(2.0) * (h0 * h0) * num;
        const double dfxn = -(hu0 / (h0 * h0)) - num / sqrt(2.0) + g * (h0 - h1) / deno;
        h0 = h0 - fxn / dfxn;
    }
    return h0;
}

}  // namespace gf
