// Host-side physics scalars of one energy step (per material).
//
// Restates the scalar physics the reference evaluates in host Python once per
// energy step: PCHIP stopping power (physics/stopping_power.py:141-228),
// linear / PCHIP coefficient interpolation (scattering.py:65-116,
// nuclear.py:94-138, coefficient_interpolation.py:51-92), Rossi transport
// coefficient (scattering.py:118-153), two-point Gauss-Legendre optical depths
// (scattering.py:267-306, nuclear.py:202-233), the integrating-factor
// exponentials (energy.py:273-310) and the Gaussian inflow spectrum
// (boundary.py:82-114).  All in IEEE double, in the reference's operation
// order.
#pragma once

#include <string>
#include <vector>

namespace m1 {

struct Table {
    std::vector<double> E, v, slope;  // slope empty when not provided
    bool empty() const { return E.empty(); }
};

struct MaterialData {
    Table sp, tr, nu;
    double density = 0.0;
    double radiation_length = 0.0;
};

// Thrown for the reference's ValueError situations (energy outside a table,
// zero energy spread, non-positive dE, ...).
struct ValueError {
    std::string msg;
};

double stopping_power(const MaterialData& m, double energy);
double transport_coefficient(const MaterialData& m, int model, int interp, double energy);
double removal_coefficient(const MaterialData& m, int interp, double energy);
double scattering_depth(const MaterialData& m, int model, int interp, double energy, double dE);
double nuclear_depth(const MaterialData& m, int interp, double energy, double dE);
// N(E; E_beam, sigma_E) of the inflow spectrum.
double energy_density(double energy, double beam_energy, double energy_spread);

}  // namespace m1
