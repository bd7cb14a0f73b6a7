// Device-side energy stepping of the batched march (m1_batch_steps.h).
//
// Per energy step of the batch:
//   steps_scalars_kernel  (before the boundary and stage launches): the CFL
//       step of every active beamlet from the CFL kernels' maxima
//       (energy.py:640-660, with the zero-state speeds of the box's vacuum
//       tiles and of the phantom outside the box folded in), then the step's
//       scalars -- S(E) at the five stage energies, the Gauss-Legendre
//       optical depths over dE/3, dE/2, 2dE/3 and dE, the integrating
//       factors, the inflow spectrum (energy.py:234-310, boundary.py:82-114)
//       -- in double and their float copies for the stage kernels; on a
//       judged step, the checkpoint each judged beamlet saves;
//   steps_advance_kernel  (after the stages): E -= dE, the judge's accept or
//       roll-back, the box flags of the next step and the CFL scalars at the
//       new energy.
// One workgroup per beamlet; its lanes evaluate the table lookups of the
// different materials and energies in parallel.
#include "m1_batch_steps.h"

#include <hip/hip_runtime.h>

#include <cstring>

namespace m1 {
namespace {

constexpr int SCALAR_THREADS = 256;   // MAXMAT materials x 32 lookups
constexpr int LOOKUPS = 32;
constexpr int ADVANCE_THREADS = 64;

__device__ inline double bits_to_double(unsigned long long u)
{
    return __longlong_as_double(static_cast<long long>(u));
}

// z = 0 tiles of a box stay evaluated while the inflow spectrum, anywhere in
// the next two maximal steps, exceeds the threshold relative to its peak.
__device__ inline int inflow_on(const BeamletConst& c, double max_step, double threshold,
                                double energy_next)
{
    const double lo = energy_next - 2.0 * max_step;
    const double e = fmin(fmax(c.energy, lo), energy_next);
    const double d = (e - c.energy) / c.energy_spread;
    return exp(-0.5 * d * d) > threshold ? 1 : 0;
}

// CFL scalars at `energy`: S(E) of every material, the inflow spectrum and
// the normalisation (fill_cfl + set_norm of the host code).
__device__ inline void cfl_scalars(const DevStepArgs& a, int b, double energy, int tid)
{
    StepCoef& c = a.cfl_coef[b];
    // a quasi-steady material sets no CFL bound (its speeds count as 0)
    if (tid < a.n_mat)
        c.S0[tid] = (a.qs_mask >> tid & 1u)
                        ? HUGE_VAL
                        : dev::stopping_power(a.mats[tid], energy, &a.ctr->err, tid, b);
    if (tid == 0) {
        const BeamletConst& q = a.bc[b];
        c.ed[0] = dev::energy_density(energy, q.energy, q.energy_spread);
        c.rf = exp(-(q.angular_spread * q.angular_spread));
        c.alpha = q.alpha;
        c.inv_alpha = 1.0 / q.alpha;
        c.eig_eps = 1.0e-10 / q.alpha;
        c.weno_eps = 1.0e-12 / (q.alpha * q.alpha);
    }
}

__global__ void __launch_bounds__(ADVANCE_THREADS) steps_init_kernel(DevStepArgs a)
{
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    __shared__ int s_on;
    const BeamletConst& q = a.bc[b];
    if (tid == 0) {
        BeamletState st{};
        st.E = q.E0;
        st.dE = 0.0;
        st.acc = -1;
        st.rollback_to = -1;
        a.st[b] = st;
        const int on = q.E0 > a.emin ? 1 : 0;
        a.step[b] = BoxStep{on, 1, a.use_model ? 1 : 0, float(q.E0), 0};
        if (on)
            atomicAdd(&a.ctr->n_active, 1);
        s_on = on;
    }
    __syncthreads();
    if (s_on)
        cfl_scalars(a, b, q.E0, tid);
}

__global__ void __launch_bounds__(SCALAR_THREADS) steps_scalars_kernel(DevStepArgs a)
{
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int nb = a.n_boxes;
    if (b == 0 && tid == 0) {
        // status of this step for the host, then this step's counters
        const int iter = a.ctr->iter;
        const int n = *a.count;
        a.ctr->tiles_evaluated += static_cast<unsigned long long>(n);
        if (a.ring) {
            StepStatus s{iter, a.ctr->n_active, n, a.ctr->err.code};
            a.ring[iter % a.ring_len] = s;
            __threadfence_system();
        }
        a.ctr->n_active = 0;
    }
    const BoxStep bs = a.step[b];
    if (!bs.active) {
        if (tid == 0 && a.sel) {
            a.sel[b] = 0;
            a.sel[nb + b] = 0;
        }
        return;  // uniform over the workgroup: no barrier has run
    }
    __shared__ double s_E, s_dE;
    __shared__ double val[MAXMAT][LOOKUPS];
    __shared__ double ts[MAXMAT][4], tn[MAXMAT][4];
    __shared__ StepCoef sc;
    __shared__ StepCoef32 sf;
    const BeamletConst& q = a.bc[b];
    if (tid == 0) {
        BeamletState& st = a.st[b];
        double mx = bits_to_double(a.red[b]);
        if (!a.vacuum_aware) {
            // cells at the zero state -- the box's tiles outside the live set
            // and, in the whole-phantom bound, the phantom outside the box:
            // isotropic speeds (wave_speeds.py:37-50) in the kernels' float
            // closure
            const double iso = a.wave == M1_WAVE_UNIT ? 1.0 : double(0.57735026918962576f);
            const unsigned long long zero_mats =
                a.box_mats[b] | (a.cfl_box_only ? 0ull : q.outside_mats);
            for (int m = 0; m < a.n_mat; ++m)
                if (zero_mats >> m & 1ull) {
                    const double s0 = a.cfl_coef[b].S0[m];
                    const double v = iso / (s0 * a.dx) + iso / (s0 * a.dy) + iso / (s0 * a.dz);
                    if (v > mx)
                        mx = v;
                }
        }
        double d = a.max_step;
        if (!dev::cfl_energy_step(a.cfl, a.max_step, mx, &d))
            dev::raise(&a.ctr->err, DEVERR_NONFINITE, 0, 0, b, mx);
        if (st.E - a.emin < d)
            d = st.E - a.emin;
        st.dE = d;
        s_E = st.E;
        s_dE = d;
        atomicAdd(&a.ctr->tiles_full, static_cast<unsigned long long>(q.tiles));
        if (a.sel) {
            uint8_t judge = 0, save = 0;
            if (a.judge_iter && bs.model_on) {
                st.pend = st.acc == 0 ? 1 : 0;
                st.ckE[st.pend] = st.E;
                st.ckS[st.pend] = st.steps;
                judge = 1;
                save = uint8_t(1 + st.pend);
            }
            a.sel[b] = judge;
            a.sel[nb + b] = save;
        }
    }
    {
        // zero the scalars (unused material slots stay zero)
        double* p = reinterpret_cast<double*>(&sc);
        for (int k = tid; k < int(sizeof(StepCoef) / sizeof(double)); k += SCALAR_THREADS)
            p[k] = 0.0;
        uint32_t* r = reinterpret_cast<uint32_t*>(&sf);
        for (int k = tid; k < int(sizeof(StepCoef32) / sizeof(uint32_t)); k += SCALAR_THREADS)
            r[k] = 0u;
    }
    __syncthreads();
    const double E = s_E, dE = s_dE;
    const double ds[4] = {dE / 3.0, 0.5 * dE, 2.0 * dE / 3.0, dE};
    const bool scat = a.scattering != M1_SCATTERING_OFF;
    {
        // one table lookup per lane: S at the five stage energies, then per
        // depth interval and Gauss node the transport coefficient, the
        // removal coefficient and S
        const int m = tid / LOOKUPS, k = tid % LOOKUPS;
        if (m < a.n_mat) {
            const DevMaterial& md = a.mats[m];
            DevError* err = &a.ctr->err;
            double v = 0.0;
            if (k < 5) {
                const double e13 = E - dE / 3.0, e23 = E - 2.0 * dE / 3.0, e1 = E - dE,
                             e2 = E - 0.5 * dE;
                const double es[5] = {E, e1, e2, e13, e23};
                v = dev::stopping_power(md, es[k], err, m, b);
            } else if (k < 29 && a.ifac) {
                const int j = (k - 5) / 6, r = (k - 5) % 6;
                const int p = r / 3, w = r % 3;
                if (ds[j] > 0.0) {
                    double lo, hi, hw;
                    dev::gauss_nodes(E, ds[j], lo, hi, hw);
                    const double x = p ? hi : lo;
                    if (w == 0 && scat)
                        v = dev::transport_coefficient(md, a.scattering, a.interp, x, err, m, b);
                    else if (w == 1 && a.nuclear)
                        v = dev::coefficient(md.nu, a.interp, x, err, m, DEVTAB_NU, b);
                    else if (w == 2 && (scat || a.nuclear))
                        v = dev::stopping_power(md, x, err, m, b);
                }
            }
            val[m][k] = v;
        }
    }
    __syncthreads();
    if (tid < MAXMAT * 4) {
        // optical depths (scattering_depth, nuclear_depth: hw (a + b))
        const int m = tid / 4, j = tid % 4;
        double s = 0.0, n = 0.0;
        if (m < a.n_mat && a.ifac && ds[j] > 0.0) {
            double lo, hi, hw;
            dev::gauss_nodes(E, ds[j], lo, hi, hw);
            const double* v = &val[m][5 + 6 * j];
            if (scat)
                s = hw * (v[0] / v[2] + v[3] / v[5]);
            if (a.nuclear)
                n = hw * (v[1] / v[2] + v[4] / v[5]);
        }
        ts[m][j] = s;
        tn[m][j] = n;
    }
    __syncthreads();
    if (tid < a.n_mat) {
        const int m = tid;
        sc.S0[m] = val[m][0];
        sc.S1[m] = val[m][1];
        sc.S2[m] = val[m][2];
        sc.S13[m] = val[m][3];
        sc.S23[m] = val[m][4];
        sc.rho[m] = a.mats[m].density;
        sc.hn[m] = sc.hs[m] = 1.0;
        // exp(+-min(tau_n, 700)), exp(+-min(tau_n + tau_s, 700))
        auto fac = [&](int j, bool inverse, double* fp, double* fj) {
            const double sign = inverse ? 1.0 : -1.0;
            const double x = fmin(tn[m][j], 700.0);
            const double y = fmin(tn[m][j] + ts[m][j], 700.0);
            *fp = exp(sign * x);
            *fj = exp(sign * y);
        };
        if (a.ifac) {
            fac(3, false, &sc.f1p[m], &sc.f1j[m]);
            fac(1, false, &sc.f2p[m], &sc.f2j[m]);
            fac(3, true, &sc.i1p[m], &sc.i1j[m]);
            fac(1, true, &sc.i2p[m], &sc.i2j[m]);
            fac(0, true, &sc.i13p[m], &sc.i13j[m]);
            fac(2, true, &sc.i23p[m], &sc.i23j[m]);
        } else {
            sc.f1p[m] = sc.f1j[m] = sc.f2p[m] = sc.f2j[m] = 1.0;
            sc.i1p[m] = sc.i1j[m] = sc.i2p[m] = sc.i2j[m] = 1.0;
            sc.i13p[m] = sc.i13j[m] = sc.i23p[m] = sc.i23j[m] = 1.0;
        }
        const double alpha = q.alpha;
        sf.S0[m] = float(sc.S0[m]);
        sf.S1[m] = float(sc.S1[m]);
        sf.S2[m] = float(sc.S2[m]);
        sf.g1p[m] = float(sc.f1p[m] / sc.S1[m]);
        sf.g1j[m] = float(sc.f1j[m] / sc.S1[m]);
        sf.g2p[m] = float(sc.f2p[m] / sc.S2[m]);
        sf.g2j[m] = float(sc.f2j[m] / sc.S2[m]);
        sf.i1p[m] = float(sc.i1p[m]);
        sf.i1j[m] = float(sc.i1j[m]);
        sf.i2p[m] = float(sc.i2p[m]);
        sf.i2j[m] = float(sc.i2j[m]);
        sf.n1p[m] = sf.g1p[m];
        sf.n1j[m] = sf.g1j[m];
        sf.dose_scale[m] = float(alpha * MEV_TO_JOULE / sc.rho[m]);
        sf.dEm[m] = (a.qs_mask >> m & 1u) ? 0.f : float(dE);
    }
    if (tid == MAXMAT * 4) {
        const double e13 = E - dE / 3.0, e23 = E - 2.0 * dE / 3.0, e1 = E - dE;
        sc.dE = dE;
        sc.ed[0] = dev::energy_density(E, q.energy, q.energy_spread);
        sc.ed[1] = dev::energy_density(e13, q.energy, q.energy_spread);
        sc.ed[2] = dev::energy_density(e23, q.energy, q.energy_spread);
        sc.ed[3] = dev::energy_density(e1, q.energy, q.energy_spread);
        sc.rf = exp(-(q.angular_spread * q.angular_spread));
        sc.alpha = q.alpha;
        sc.inv_alpha = 1.0 / q.alpha;
        sc.eig_eps = 1.0e-10 / q.alpha;
        sc.weno_eps = 1.0e-12 / (q.alpha * q.alpha);
        sf.dE = float(dE);
        sf.weno_eps = float(fmax(sc.weno_eps, 1.0e-37));
        sf.eig_eps = sc.eig_eps;
    }
    __syncthreads();
    {
        const double* p = reinterpret_cast<const double*>(&sc);
        double* o = reinterpret_cast<double*>(a.coef + b);
        for (int k = tid; k < int(sizeof(StepCoef) / sizeof(double)); k += SCALAR_THREADS)
            o[k] = p[k];
        const uint32_t* r = reinterpret_cast<const uint32_t*>(&sf);
        uint32_t* w = reinterpret_cast<uint32_t*>(a.coef32 + b);
        for (int k = tid; k < int(sizeof(StepCoef32) / sizeof(uint32_t)); k += SCALAR_THREADS)
            w[k] = r[k];
    }
}

__global__ void __launch_bounds__(ADVANCE_THREADS) steps_advance_kernel(DevStepArgs a)
{
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int nb = a.n_boxes;
    __shared__ int s_on;
    __shared__ double s_E;
    if (tid == 0) {
        if (b == 0) {
            const int next = a.ctr->iter + 1;
            a.ctr->iter = next;
            if (a.gate)
                *a.gate = a.use_model && a.infer_every > 0 && next % a.infer_every == 0 ? 1 : 0;
            if (a.judge_iter && a.count_full)
                a.ctr->tiles_evaluated += static_cast<unsigned long long>(*a.count_full);
        }
        BoxStep bs = a.step[b];
        bs.full_set = 0;
        if (a.judge_iter && a.sel)
            a.sel[2 * nb + b] = 0;
        int on = 0;
        if (bs.active) {
            BeamletState st = a.st[b];
            const BeamletConst& q = a.bc[b];
            st.E -= st.dE;
            st.steps += 1;
            if (a.judge_iter && a.sel && a.sel[b]) {
                // accepted: the pending checkpoint becomes the last good
                // one; rejected: back to the last good checkpoint, on every
                // tile for one step, without the network after
                const float df = __uint_as_float(a.jdiff[b]);
                const float fl = __uint_as_float(a.jfull[b]);
                atomicAdd(&a.ctr->judged, 1ull);
                if (!(df <= a.judge_tol * fl) && !(fl == 0.f && df == 0.f)) {
                    st.rejected = 1;
                    atomicAdd(&a.ctr->rejections, 1ull);
                    a.sel[2 * nb + b] = uint8_t(st.acc < 0 ? 3 : 1 + st.acc);
                    st.E = st.acc < 0 ? q.E0 : st.ckE[st.acc];
                    st.steps = st.acc < 0 ? 0 : st.ckS[st.acc];
                    st.rollback_to = st.steps;
                    bs.full_set = 1;
                } else {
                    st.acc = st.pend;
                }
            }
            on = st.E > a.emin ? 1 : 0;
            bs.active = on;
            bs.inflow_on = on ? inflow_on(q, a.max_step, a.tile_threshold, st.E) : 0;
            bs.model_on = a.use_model && !st.rejected ? 1 : 0;
            bs.energy = float(st.E);
            a.st[b] = st;
            if (on)
                atomicAdd(&a.ctr->n_active, 1);
            s_E = st.E;
        }
        a.step[b] = bs;
        s_on = on;
    }
    __syncthreads();
    if (s_on)
        cfl_scalars(a, b, s_E, tid);
}

__global__ void physics_probe_kernel(const DevMaterial* mats, int m, int kind, int model,
                                     int interp, const double* energies, double dE, int n,
                                     double* out, DevError* err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const DevMaterial& md = mats[m];
    const double e = energies[i];
    double r = 0.0;
    if (kind == 0) {
        r = dev::stopping_power(md, e, err, m, i);
    } else if (kind == 1) {
        r = dev::transport_coefficient(md, model, interp, e, err, m, i);
    } else if (kind == 2) {
        r = dev::coefficient(md.nu, interp, e, err, m, DEVTAB_NU, i);
    } else {
        double lo, hi, hw;
        dev::gauss_nodes(e, dE, lo, hi, hw);
        if (kind == 3) {
            if (model != M1_SCATTERING_OFF)
                r = hw * (dev::transport_coefficient(md, model, interp, lo, err, m, i) /
                              dev::stopping_power(md, lo, err, m, i) +
                          dev::transport_coefficient(md, model, interp, hi, err, m, i) /
                              dev::stopping_power(md, hi, err, m, i));
        } else {
            r = hw * (dev::coefficient(md.nu, interp, lo, err, m, DEVTAB_NU, i) /
                          dev::stopping_power(md, lo, err, m, i) +
                      dev::coefficient(md.nu, interp, hi, err, m, DEVTAB_NU, i) /
                          dev::stopping_power(md, hi, err, m, i));
        }
    }
    out[i] = r;
}

}  // namespace

hipError_t launch_steps_init(const DevStepArgs& a, hipStream_t s)
{
    if (a.n_boxes <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(steps_init_kernel, dim3(unsigned(a.n_boxes)), dim3(ADVANCE_THREADS), 0, s,
                       a);
    return hipGetLastError();
}

hipError_t launch_steps_scalars(const DevStepArgs& a, hipStream_t s)
{
    if (a.n_boxes <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(steps_scalars_kernel, dim3(unsigned(a.n_boxes)), dim3(SCALAR_THREADS), 0,
                       s, a);
    return hipGetLastError();
}

hipError_t launch_steps_advance(const DevStepArgs& a, hipStream_t s)
{
    if (a.n_boxes <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(steps_advance_kernel, dim3(unsigned(a.n_boxes)), dim3(ADVANCE_THREADS), 0,
                       s, a);
    return hipGetLastError();
}

hipError_t launch_physics_probe(const DevMaterial* mats, int m, int kind, int model, int interp,
                                const double* energies, double dE, int n, double* out,
                                DevError* err, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(physics_probe_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                       mats, m, kind, model, interp, energies, dE, n, out, err);
    return hipGetLastError();
}

}  // namespace m1
