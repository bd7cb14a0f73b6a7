// Device form of the per-step physics scalars (m1_physics.cpp, m1_host.h),
// so that a batched energy step is set up without a host round trip.
//
// Same arithmetic as the host code, operation for operation (the translation
// units that include this header are compiled with -ffp-contract=off):
// PCHIP stopping power (physics/stopping_power.py:141-228), linear / PCHIP
// coefficient interpolation (scattering.py:65-116, nuclear.py:94-138,
// coefficient_interpolation.py:51-92), Rossi transport coefficient
// (scattering.py:118-153), two-point Gauss-Legendre optical depths
// (scattering.py:267-306, nuclear.py:202-233), integrating-factor
// exponentials (energy.py:273-310), the Gaussian inflow spectrum
// (boundary.py:82-114) and the CFL energy step (energy.py:640-660).  Sums,
// products, quotients and square roots round exactly as on the host; exp is
// the device library's (within an ulp of the host's, not always equal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "m1_device.h"
#include "m1march.h"

namespace m1 {

struct DevTable {
    const double* E;
    const double* v;
    const double* slope;  // null when the table has no nodal slopes
    int n;
    int pad;
};

struct DevMaterial {
    DevTable sp, tr, nu;
    double density;
    double radiation_length;
};

// First error a device step met (the host raises it): an energy outside a
// table, a non-finite CFL maximum, PCHIP without slopes.
enum DevErrorCode : int { DEVERR_NONE = 0, DEVERR_RANGE = 1, DEVERR_NONFINITE = 2,
                          DEVERR_SLOPES = 3 };
enum DevTableKind : int { DEVTAB_SP = 0, DEVTAB_TR = 1, DEVTAB_NU = 2 };

struct DevError {
    int code;
    int material;
    int table;
    int beamlet;
    double energy;
};

namespace dev {

__device__ inline void raise(DevError* e, int code, int material, int table, int beamlet,
                             double energy)
{
    if (atomicCAS(&e->code, 0, -1) == 0) {
        e->material = material;
        e->table = table;
        e->beamlet = beamlet;
        e->energy = energy;
        __threadfence();
        atomicExch(&e->code, code);
    }
}

// numpy.isclose(a, b) with its default tolerances.
__device__ inline bool isclose(double a, double b)
{
    return fabs(a - b) <= (1e-08 + 1e-05 * fabs(b));
}

// Index of the first node above x (std::upper_bound).
__device__ inline int upper_bound(const DevTable& t, double x)
{
    int first = 0, count = t.n;
    while (count > 0) {
        const int step = count / 2;
        const int it = first + step;
        if (!(x < t.E[it])) {
            first = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    return first;
}

// Index i with E[i] <= x < E[i+1], clamped to [0, n-2].
__device__ inline int interval(const DevTable& t, double x)
{
    int i = upper_bound(t, x);
    i = i == 0 ? 0 : i - 1;
    return i < t.n - 2 ? i : t.n - 2;
}

// Energy clamped into the table; records a range error as the host throws it.
__device__ inline double in_table(const DevTable& t, double energy, DevError* err, int material,
                                  int table, int beamlet)
{
    const double lo = t.E[0], hi = t.E[t.n - 1];
    if ((energy < lo && !isclose(energy, lo)) || (energy > hi && !isclose(energy, hi)))
        raise(err, DEVERR_RANGE, material, table, beamlet, energy);
    return fmin(fmax(energy, lo), hi);
}

__device__ inline double hermite(const DevTable& t, double energy)
{
    if (energy >= t.E[t.n - 1])
        return t.v[t.n - 1];
    const int i = interval(t, energy);
    const double h = t.E[i + 1] - t.E[i];
    const double s = (energy - t.E[i]) / h;
    const double s2 = s * s;
    const double s3 = s2 * s;
    return (2.0 * s3 - 3.0 * s2 + 1.0) * t.v[i] + (s3 - 2.0 * s2 + s) * h * t.slope[i] +
           (-2.0 * s3 + 3.0 * s2) * t.v[i + 1] + (s3 - s2) * h * t.slope[i + 1];
}

// numpy.interp at one abscissa inside the table.
__device__ inline double linear(const DevTable& t, double x)
{
    if (x >= t.E[t.n - 1])
        return t.v[t.n - 1];
    const int j = interval(t, x);
    if (t.E[j] == x)
        return t.v[j];
    const double slope = (t.v[j + 1] - t.v[j]) / (t.E[j + 1] - t.E[j]);
    double r = slope * (x - t.E[j]) + t.v[j];
    if (isnan(r)) {
        r = slope * (x - t.E[j + 1]) + t.v[j + 1];
        if (isnan(r) && t.v[j] == t.v[j + 1])
            r = t.v[j];
    }
    return r;
}

__device__ inline double stopping_power(const DevMaterial& m, double energy, DevError* err,
                                        int material, int beamlet)
{
    energy = in_table(m.sp, energy, err, material, DEVTAB_SP, beamlet);
    return hermite(m.sp, energy);
}

__device__ inline double coefficient(const DevTable& t, int interp, double energy, DevError* err,
                                     int material, int table, int beamlet)
{
    energy = in_table(t, energy, err, material, table, beamlet);
    if (interp == M1_INTERP_LINEAR)
        return linear(t, energy);
    if (!t.slope) {
        raise(err, DEVERR_SLOPES, material, table, beamlet, energy);
        return 0.0;
    }
    return hermite(t, energy);
}

__device__ inline double transport_coefficient(const DevMaterial& m, int model, int interp,
                                               double energy, DevError* err, int material,
                                               int beamlet)
{
    if (model == M1_SCATTERING_WENTZEL)
        return coefficient(m.tr, interp, energy, err, material, DEVTAB_TR, beamlet);
    if (model == M1_SCATTERING_ROSSI) {
        const double e = fmax(energy, 1.0e-8);
        const double mp = 938.27208816;
        const double pv = e * (e + 2.0 * mp) / (e + mp);
        const double r = 15.0 / pv;
        return r * r / m.radiation_length;
    }
    return 0.0;
}

// Gauss-Legendre nodes of [energy - dE, energy]: mid -/+ off.
__device__ inline void gauss_nodes(double energy, double dE, double& lo, double& hi, double& hw)
{
    const double mid = energy - 0.5 * dE;
    hw = 0.5 * dE;
    const double off = hw / sqrt(3.0);
    lo = mid - off;
    hi = mid + off;
}

// N(E; E_beam, sigma_E) of the inflow spectrum (sigma_E != 0, checked when
// the problem is bound).
__device__ inline double energy_density(double energy, double beam_energy, double s)
{
    const double d = energy - beam_energy;
    return exp(-(d * d) / (2.0 * (s * s))) / (sqrt(2.0 * M_PI) * s);
}

// Energy step of the CFL maximum `mx` (energy.py:651-660); false when mx is
// not finite.
__device__ inline bool cfl_energy_step(double cfl, double max_step, double mx, double* dE)
{
    if (!isfinite(mx))
        return false;
    const double d = cfl / mx;
    *dE = d <= max_step ? d : max_step;
    return true;
}

}  // namespace dev
}  // namespace m1
