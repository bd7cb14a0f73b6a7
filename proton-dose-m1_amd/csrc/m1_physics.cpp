// Host-side physics scalars; see m1_physics.h for the reference functions.
#include "m1_physics.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "m1march.h"

namespace m1 {
namespace {

// numpy.isclose(a, b) with its default tolerances.
bool isclose(double a, double b) { return std::fabs(a - b) <= (1e-08 + 1e-05 * std::fabs(b)); }

double in_table(const Table& t, double energy, const char* what)
{
    const double lo = t.E.front(), hi = t.E.back();
    if ((energy < lo && !isclose(energy, lo)) || (energy > hi && !isclose(energy, hi))) {
        char buf[256];
        std::snprintf(buf, sizeof buf, "Energy %.17g MeV outside %s range [%.17g, %.17g] MeV.",
                      energy, what, lo, hi);
        throw ValueError{buf};
    }
    return std::min(std::max(energy, lo), hi);
}

// Index i with E[i] <= x < E[i+1], clamped to [0, n-2].
size_t interval(const Table& t, double x)
{
    const size_t n = t.E.size();
    size_t i = size_t(std::upper_bound(t.E.begin(), t.E.end(), x) - t.E.begin());
    i = i == 0 ? 0 : i - 1;
    return std::min(i, n - 2);
}

// Cubic Hermite on [E_i, E_i+1] with nodal slopes; energy already clamped.
double hermite(const Table& t, double energy)
{
    if (energy >= t.E.back())
        return t.v.back();
    const size_t i = interval(t, energy);
    const double h = t.E[i + 1] - t.E[i];
    const double s = (energy - t.E[i]) / h;
    const double s2 = s * s;
    const double s3 = s2 * s;
    return (2.0 * s3 - 3.0 * s2 + 1.0) * t.v[i] + (s3 - 2.0 * s2 + s) * h * t.slope[i] +
           (-2.0 * s3 + 3.0 * s2) * t.v[i + 1] + (s3 - s2) * h * t.slope[i + 1];
}

// numpy.interp at one abscissa inside the table.
double linear(const Table& t, double x)
{
    const size_t n = t.E.size();
    if (x >= t.E[n - 1])
        return t.v[n - 1];
    const size_t j = interval(t, x);
    if (t.E[j] == x)
        return t.v[j];
    const double slope = (t.v[j + 1] - t.v[j]) / (t.E[j + 1] - t.E[j]);
    double r = slope * (x - t.E[j]) + t.v[j];
    if (std::isnan(r)) {
        r = slope * (x - t.E[j + 1]) + t.v[j + 1];
        if (std::isnan(r) && t.v[j] == t.v[j + 1])
            r = t.v[j];
    }
    return r;
}

double coefficient(const Table& t, int interp, double energy, const char* what)
{
    energy = in_table(t, energy, what);
    if (interp == M1_INTERP_LINEAR)
        return linear(t, energy);
    if (t.slope.empty())
        throw ValueError{"PCHIP interpolation needs nodal slopes"};
    return hermite(t, energy);
}

}  // namespace

double stopping_power(const MaterialData& m, double energy)
{
    energy = in_table(m.sp, energy, "stopping-power table");
    return hermite(m.sp, energy);
}

double transport_coefficient(const MaterialData& m, int model, int interp, double energy)
{
    if (model == M1_SCATTERING_WENTZEL)
        return coefficient(m.tr, interp, energy, "Wentzel table");
    if (model == M1_SCATTERING_ROSSI) {
        const double e = std::max(energy, 1.0e-8);
        const double mp = 938.27208816;
        const double pv = e * (e + 2.0 * mp) / (e + mp);
        const double r = 15.0 / pv;
        return r * r / m.radiation_length;
    }
    return 0.0;
}

double removal_coefficient(const MaterialData& m, int interp, double energy)
{
    return coefficient(m.nu, interp, energy, "nuclear table");
}

double scattering_depth(const MaterialData& m, int model, int interp, double energy, double dE)
{
    if (model == M1_SCATTERING_OFF)
        return 0.0;
    const double mid = energy - 0.5 * dE;
    const double hw = 0.5 * dE;
    const double off = hw / std::sqrt(3.0);
    const double a = transport_coefficient(m, model, interp, mid - off) / stopping_power(m, mid - off);
    const double b = transport_coefficient(m, model, interp, mid + off) / stopping_power(m, mid + off);
    return hw * (a + b);
}

double nuclear_depth(const MaterialData& m, int interp, double energy, double dE)
{
    if (dE <= 0.0)
        throw ValueError{"dE must be positive."};
    const double mid = energy - 0.5 * dE;
    const double hw = 0.5 * dE;
    const double off = hw / std::sqrt(3.0);
    const double a = removal_coefficient(m, interp, mid - off) / stopping_power(m, mid - off);
    const double b = removal_coefficient(m, interp, mid + off) / stopping_power(m, mid + off);
    return hw * (a + b);
}

double energy_density(double energy, double beam_energy, double s)
{
    if (s == 0.0)
        throw ValueError{"Zero energy spread is not supported by the Gaussian inflow."};
    const double d = energy - beam_energy;
    return std::exp(-(d * d) / (2.0 * (s * s))) / (std::sqrt(2.0 * M_PI) * s);
}

}  // namespace m1
