// Host-side pieces shared by the single-box march (m1_march.hip) and the
// batched-beamlet march (m1_batch_march.hip): error types, numerics
// validation, material tables, the per-step physics scalars and the
// HIP-event kernel profiler.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "m1_device.h"
#include "m1_physics.h"
#include "m1march.h"

namespace m1 {

struct HipError {
    std::string msg;
};

struct NotImplemented {
    std::string msg;
};

struct NonFinite {
    std::string msg;
};

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
}

template <class T>
void dev_alloc(T** p, size_t count, const char* what)
{
    hip_check(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)), what);
}

template <class T>
void host_alloc(T** p, size_t count, const char* what)
{
    hip_check(hipHostMalloc(reinterpret_cast<void**>(p), count * sizeof(T)), what);
}

// Status code and message of a caught exception (the ABI's error contract).
template <class F>
int run_guarded(int device, std::string& err, F&& f)
{
    try {
        hip_check(hipSetDevice(device), "hipSetDevice");
        f();
        err.clear();
        return M1_OK;
    } catch (const ValueError& e) {
        err = e.msg;
        return M1_ERR_VALUE;
    } catch (const NotImplemented& e) {
        err = e.msg;
        return M1_ERR_NOT_IMPLEMENTED;
    } catch (const NonFinite& e) {
        err = e.msg;
        return M1_ERR_NONFINITE;
    } catch (const HipError& e) {
        err = e.msg;
        return M1_ERR_HIP;
    } catch (const std::exception& e) {
        err = e.what();
        return M1_ERR_HIP;
    }
}

// ------------------------------------------------------------ numerics
struct Num {
    int scheme, wave, scattering, nuclear, interp, precision;
    double max_step, cfl, emin, window, tile_threshold;
    int vacuum_aware_cfl;
    int batch_cfl_box_only;
};

// Validation in the order and wording of config.py:53-89 for the fields the
// march reads.
inline Num read_numerics(const m1_numerics* n)
{
    if (!n)
        throw ValueError{"numerics must not be NULL"};
    Num r{n->energy_scheme, n->wave_speeds, n->scattering, n->nuclear,
          n->coefficient_interpolation, n->precision, n->max_energy_step, n->cfl,
          n->energy_min, n->energy_window_sigma, n->tile_threshold, n->vacuum_aware_cfl != 0,
          n->batch_cfl_box_only != 0};
    if (r.scheme < M1_FORWARD_EULER || r.scheme > M1_IF_SSPRK3)
        throw NotImplemented{"Energy scheme has not been migrated yet."};
    if (r.wave != M1_WAVE_EXACT && r.wave != M1_WAVE_UNIT)
        throw ValueError{"Unknown wave-speed mode."};
    if (r.scattering < M1_SCATTERING_OFF || r.scattering > M1_SCATTERING_WENTZEL)
        throw ValueError{"Unknown scattering model."};
    if (r.interp != M1_INTERP_LINEAR && r.interp != M1_INTERP_PCHIP)
        throw ValueError{"Unknown coefficient interpolation."};
    if (r.precision < M1_FP64 || r.precision > M1_FP32_BULK_F64_CLOSURE)
        throw ValueError{"Unknown precision."};
    if (!(std::isfinite(r.max_step) && std::isfinite(r.cfl) && std::isfinite(r.emin) &&
          std::isfinite(r.window)))
        throw ValueError{"Numerical parameters must be finite."};
    if (r.max_step <= 0)
        throw ValueError{"max_energy_step must be positive."};
    if (!(r.cfl > 0 && r.cfl <= 1))
        throw ValueError{"cfl must lie in (0, 1]."};
    if (r.emin < 0)
        throw ValueError{"energy_min cannot be negative."};
    if (!(r.tile_threshold >= 0.0 && r.tile_threshold < 1.0))
        throw ValueError{"tile_threshold must lie in [0, 1)."};
    return r;
}

inline bool has_sources(const Num& n) { return n.scattering != M1_SCATTERING_OFF || n.nuclear; }

// The march runs the float-state kernels.
inline bool is_f32(const Num& n)
{
    return n.precision == M1_FP32_BULK || n.precision == M1_FP32_BULK_F64_CLOSURE;
}

// Material tables of the ABI, checked as stopping_power.py:73-91 checks them.
inline MaterialData read_material(const m1_material& m, int idx)
{
    auto table = [&](const m1_table& t, bool need_slope, const char* what) {
        Table r;
        if (t.n < 2 || !t.energy || !t.value)
            throw ValueError{std::string(what) + " table of material " + std::to_string(idx) +
                             " must contain at least two points."};
        r.E.assign(t.energy, t.energy + t.n);
        r.v.assign(t.value, t.value + t.n);
        for (int i = 1; i < t.n; ++i)
            if (!(r.E[i] > r.E[i - 1]))
                throw ValueError{std::string(what) + " energies must be strictly increasing."};
        if (t.slope)
            r.slope.assign(t.slope, t.slope + t.n);
        else if (need_slope)
            throw ValueError{std::string(what) + " table needs PCHIP slopes."};
        return r;
    };
    MaterialData d;
    d.sp = table(m.stopping_power, true, "Stopping-power");
    for (double v : d.sp.v)
        if (!(std::isfinite(v) && v > 0.0))
            throw ValueError{"Stopping-power values must be finite and positive."};
    d.tr = table(m.transport, false, "Transport");
    d.nu = table(m.removal, false, "Nuclear-attenuation");
    d.density = m.density;
    d.radiation_length = m.radiation_length;
    if (!(d.density > 0.0 && std::isfinite(d.density)))
        throw ValueError{"material density must be finite and positive"};
    return d;
}

// ------------------------------------------------------- step scalars
struct BeamParams {
    double energy, energy_spread, angular_spread;
};

// Per-material optical depths over [E - d, E] (energy.py:234-270).
inline void depths(const std::vector<MaterialData>& mats, const Num& n, double energy, double d,
                   double* ts, double* tn)
{
    for (size_t m = 0; m < mats.size(); ++m) {
        ts[m] = 0.0;
        tn[m] = 0.0;
        if (d <= 0.0)
            continue;
        if (n.scattering != M1_SCATTERING_OFF)
            ts[m] = scattering_depth(mats[m], n.scattering, n.interp, energy, d);
        if (n.nuclear)
            tn[m] = nuclear_depth(mats[m], n.interp, energy, d);
    }
}

// exp(+-min(tau_n, 700)), exp(+-min(tau_n + tau_s, 700)) (energy.py:294-308).
inline void factors(size_t nm, const double* ts, const double* tn, bool inverse, double* fp,
                    double* fj)
{
    const double sign = inverse ? 1.0 : -1.0;
    for (size_t m = 0; m < nm; ++m) {
        const double a = std::min(tn[m], 700.0);
        const double b = std::min(tn[m] + ts[m], 700.0);
        fp[m] = std::exp(sign * a);
        fj[m] = std::exp(sign * b);
    }
}

// All scalars of one step at (energy, dE).
inline void fill_step(const std::vector<MaterialData>& mats, const BeamParams& bp, const Num& n,
                      double energy, double dE, bool ifac, bool strang, StepCoef& s)
{
    std::memset(&s, 0, sizeof s);
    const size_t nm = mats.size();
    s.dE = dE;
    const double e13 = energy - dE / 3.0, e23 = energy - 2.0 * dE / 3.0, e1 = energy - dE,
                 e2 = energy - 0.5 * dE;
    s.ed[0] = energy_density(energy, bp.energy, bp.energy_spread);
    s.ed[1] = energy_density(e13, bp.energy, bp.energy_spread);
    s.ed[2] = energy_density(e23, bp.energy, bp.energy_spread);
    s.ed[3] = energy_density(e1, bp.energy, bp.energy_spread);
    s.rf = std::exp(-(bp.angular_spread * bp.angular_spread));
    for (size_t m = 0; m < nm; ++m) {
        const MaterialData& md = mats[m];
        s.S0[m] = stopping_power(md, energy);
        s.S1[m] = stopping_power(md, e1);
        s.S2[m] = stopping_power(md, e2);
        s.S13[m] = stopping_power(md, e13);
        s.S23[m] = stopping_power(md, e23);
        s.rho[m] = md.density;
        s.f1p[m] = s.f1j[m] = s.f2p[m] = s.f2j[m] = 1.0;
        s.i1p[m] = s.i1j[m] = s.i2p[m] = s.i2j[m] = 1.0;
        s.i13p[m] = s.i13j[m] = s.i23p[m] = s.i23j[m] = 1.0;
        s.hn[m] = s.hs[m] = 1.0;
    }
    if (ifac) {
        double ts[4][MAXMAT], tn[4][MAXMAT];
        const double ds[4] = {dE / 3.0, 0.5 * dE, 2.0 * dE / 3.0, dE};
        for (int q = 0; q < 4; ++q)
            depths(mats, n, energy, ds[q], ts[q], tn[q]);
        factors(nm, ts[3], tn[3], false, s.f1p, s.f1j);
        factors(nm, ts[1], tn[1], false, s.f2p, s.f2j);
        factors(nm, ts[3], tn[3], true, s.i1p, s.i1j);
        factors(nm, ts[1], tn[1], true, s.i2p, s.i2j);
        factors(nm, ts[0], tn[0], true, s.i13p, s.i13j);
        factors(nm, ts[2], tn[2], true, s.i23p, s.i23j);
    }
    if (strang) {
        // energy.py:1206-1227 with the half-step factors of 1168-1191.
        for (size_t m = 0; m < nm; ++m) {
            double eta = 0.0, taun = 0.0;
            if (n.scattering != M1_SCATTERING_OFF)
                eta = scattering_depth(mats[m], n.scattering, n.interp, energy, dE);
            if (n.nuclear)
                taun = nuclear_depth(mats[m], n.interp, energy, dE);
            s.hn[m] = std::exp(-std::min(0.5 * taun, 700.0));
            s.hs[m] = std::exp(-std::min(0.5 * eta, 700.0));
        }
    }
}

// Scalars of the CFL reduction at `energy` (energy.py:539-660 reads S(E) and
// the inflow state at E).
inline void fill_cfl(const std::vector<MaterialData>& mats, const BeamParams& bp, double energy,
                     StepCoef& s)
{
    std::memset(&s, 0, sizeof s);
    for (size_t m = 0; m < mats.size(); ++m)
        s.S0[m] = stopping_power(mats[m], energy);
    s.ed[0] = energy_density(energy, bp.energy, bp.energy_spread);
    s.rf = std::exp(-(bp.angular_spread * bp.angular_spread));
}

// Peak inflow fluence max(profile) * N(E_beam): the FP32-bulk state is U / alpha.
inline double march_alpha(double profile_max, const BeamParams& bp)
{
    const double a = profile_max * energy_density(bp.energy, bp.energy, bp.energy_spread);
    return a > 0.0 && std::isfinite(a) ? a : 1.0;
}

inline void set_norm(double alpha, StepCoef& s)
{
    s.alpha = alpha;
    s.inv_alpha = 1.0 / alpha;
    s.eig_eps = 1.0e-10 / alpha;
    s.weno_eps = 1.0e-12 / (alpha * alpha);
}

// Float stage scalars from the double ones.
inline void fill_step32(size_t nm, double alpha, const StepCoef& d, StepCoef32& f)
{
    std::memset(&f, 0, sizeof f);
    f.dE = float(d.dE);
    f.weno_eps = float(std::max(d.weno_eps, 1.0e-37));
    f.eig_eps = d.eig_eps;
    for (size_t m = 0; m < nm; ++m) {
        f.S0[m] = float(d.S0[m]);
        f.S1[m] = float(d.S1[m]);
        f.S2[m] = float(d.S2[m]);
        f.g1p[m] = float(d.f1p[m] / d.S1[m]);
        f.g1j[m] = float(d.f1j[m] / d.S1[m]);
        f.g2p[m] = float(d.f2p[m] / d.S2[m]);
        f.g2j[m] = float(d.f2j[m] / d.S2[m]);
        f.i1p[m] = float(d.i1p[m]);
        f.i1j[m] = float(d.i1j[m]);
        f.i2p[m] = float(d.i2p[m]);
        f.i2j[m] = float(d.i2j[m]);
        f.n1p[m] = f.g1p[m];
        f.n1j[m] = f.g1j[m];
        f.dose_scale[m] = float(alpha * MEV_TO_JOULE / d.rho[m]);
        f.dEm[m] = f.dE;
    }
}

// Energy step of the CFL maximum `mx` (energy.py:651-660).
inline double cfl_energy_step(const Num& n, double mx)
{
    if (!std::isfinite(mx))
        throw NonFinite{"non-finite M1 state detected by the CFL reduction"};
    const double dE = n.cfl / mx;
    return dE <= n.max_step ? dE : n.max_step;
}

// ------------------------------------------------------------ profiling
enum KernelId {
    K_STAGE0, K_STAGE1, K_STAGE2, K_RHS, K_FE, K_CFL, K_BOUNDARY, K_SCALE, K_TILE_MASK,
    K_TILE_ZERO, K_DOSE, K_MODEL, K_STEP_SCALARS, K_STEP_ADVANCE, K_RELAX, K_COUNT
};
inline const char* kernel_name(int k)
{
    static const char* const names[K_COUNT] = {
        "stage0", "stage1", "stage2", "spatial_operator", "forward_euler", "cfl", "boundary",
        "source_half_step", "tile_mask", "tile_zero", "dose", "tile_skip_network",
        "step_scalars", "step_advance", "quasi_steady_relax"};
    return k >= 0 && k < K_COUNT ? names[k] : "";
}

// Per-kernel launch counts and HIP-event durations on one stream.
struct Profiler {
    bool on = false;
    struct Pending {
        int kernel;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Pending> pending;
    int64_t launches[K_COUNT] = {};
    double total_ms[K_COUNT] = {};
    int64_t tiles_evaluated = 0, tiles_full = 0;

    hipEvent_t take()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    // Fold finished event pairs into the statistics (after a stream sync).
    void collect()
    {
        for (auto& p : pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                launches[p.kernel] += 1;
                total_ms[p.kernel] += ms;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset()
    {
        tiles_evaluated = tiles_full = 0;
        for (int k = 0; k < K_COUNT; ++k) {
            launches[k] = 0;
            total_ms[k] = 0.0;
        }
    }
    void release()
    {
        for (auto& e : pool)
            (void)hipEventDestroy(e);
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        pool.clear();
        pending.clear();
    }
};

// Brackets one launch with an event pair when profiling is on.
struct Timed {
    Profiler& p;
    hipStream_t s;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(Profiler& prof, hipStream_t stream, int kernel) : p(prof), s(stream), k(kernel)
    {
        if (p.on) {
            a = p.take();
            b = p.take();
            hip_check(hipEventRecord(a, s), "hipEventRecord");
        }
    }
    ~Timed()
    {
        if (p.on) {
            (void)hipEventRecord(b, s);
            p.pending.push_back({k, a, b});
        }
    }
};

}  // namespace m1
