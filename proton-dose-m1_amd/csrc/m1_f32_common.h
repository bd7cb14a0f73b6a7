// FP32-bulk stage arithmetic shared by the single-box kernels
// (m1_kernels_f32.hip) and the batched-beamlet kernels (m1_batch_kernels.hip).
// See m1_kernels_f32.hip for the precision split.
//
// These kernels are compiled with -ffp-contract=off: every fused multiply-add
// is written out (fmaf), so that the four moments of a cell always go
// through identical roundings.  A beam whose J equals phi along an axis then
// keeps J = phi bit for bit; compiler-chosen fusion was found to differ
// between the moments and to push such a beam off the realizability margin,
// where the closure is far more diffusive across the beam.
#pragma once

#include "m1_device.h"

namespace m1 {
namespace f32 {

// Everything one workgroup needs to evaluate one tile of one box: pointers
// already offset to the box, the box grid, and the tile coordinates.
struct TileView {
    GridDims g;
    const float* src;
    float* U;
    const float* V1;
    float* out;
    float* dose_hi;
    float* dose_lo;
    float* edep;
    const uint8_t* ids;
    const float* bnd;
    const StepCoef32* coef;
    float* tile_max;   // this tile's slot (nullable); relaxation: the box's change slot
    float* tile_sums;  // this tile's sum phi, sum |J| (nullable)
    float* tile_face;  // this tile's FACE_SLOTS face slabs (nullable; needs tile_max)
    int wave_mode;
    int tx, ty, tz;    // tile coordinates (x: TX planes, y: TY rows, z: TZ cells)
    // Stages 1 and 2 (nullable): the stencil input proj(C_k V_k / S_k),
    // already formed by the previous stage, read in place of src.
    const float* win;
    // Stages 0 and 1 (nullable): this stage's output as the next stage's
    // stencil input, proj(C_k out / S_k), for its win.
    float* wout;
};

// Waves per SIMD the stage kernels are compiled for (register budget:
// 512 / waves VGPRs per lane).
#ifndef M1_STAGE_WAVES
#define M1_STAGE_WAVES 3
#endif

namespace detail {

constexpr int TY = F32_TY;
constexpr int TZ = F32_TZ;
constexpr int TX = F32_TX;
static_assert(TY * TZ == 256, "one thread per y-z cell of a tile");
static_assert(TX * TY <= 256, "the batch CFL kernel maps one thread to each z-min column");
constexpr int NT = TY * TZ;
static_assert(TY <= 64 && TZ <= 64 && NT >= 128,
              "upper tile faces: z in the first wave, y in the second");
constexpr int PY = TY + 6;
constexpr int PZ = TZ + 6;
constexpr int NHALO = 6 * TZ + 6 * TY;
static_assert(NHALO <= NT, "one y-z halo cell per thread");
constexpr double REAL_EPS = 1.0e-10;  // weno_z.py:10, realizability.py:182

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// WENO-Z5 left state (weno_z.py:13-87) in float.  The ratios tau5/(beta+eps)
// are capped at 1e18: beyond it the dominant weight is 1 to float precision
// and the cap keeps alpha_k = d_k (1 + r_k^2) finite.
__device__ __forceinline__ float weno_left(float v0, float v1, float v2, float v3, float v4,
                                           float eps)
{
    const float p0 = fmaf(v0, 1.f / 3.f, fmaf(v1, -7.f / 6.f, v2 * (11.f / 6.f)));
    const float p1 = fmaf(v2, 5.f / 6.f, fmaf(v1, -1.f / 6.f, v3 * (1.f / 3.f)));
    const float p2 = fmaf(v2, 1.f / 3.f, fmaf(v3, 5.f / 6.f, v4 * (-1.f / 6.f)));
    float a = fmaf(-2.f, v1, v0) + v2;
    float b = fmaf(3.f, v2, fmaf(-4.f, v1, v0));
    const float b0 = fmaf((13.f / 12.f) * a, a, 0.25f * (b * b));
    a = fmaf(-2.f, v2, v1) + v3;
    b = v1 - v3;
    const float b1 = fmaf((13.f / 12.f) * a, a, 0.25f * (b * b));
    a = fmaf(-2.f, v3, v2) + v4;
    b = fmaf(-4.f, v3, fmaf(3.f, v2, v4));
    const float b2 = fmaf((13.f / 12.f) * a, a, 0.25f * (b * b));
    const float tau = fabsf(b2 - b0);
    const float r0 = fminf(tau * rcpf(b0 + eps), 1e18f);
    const float r1 = fminf(tau * rcpf(b1 + eps), 1e18f);
    const float r2 = fminf(tau * rcpf(b2 + eps), 1e18f);
    const float a0 = fmaf(0.1f * r0, r0, 0.1f);
    const float a1 = fmaf(0.6f * r1, r1, 0.6f);
    const float a2 = fmaf(0.3f * r2, r2, 0.3f);
    return fmaf(a0, p0, fmaf(a1, p1, a2 * p2)) * rcpf(a0 + a1 + a2);
}

// Both WENO-Z5 states of one cell from its stencil v0..v4 (centre v2): the
// left state at its upper face, weno_left(v0..v4), and the right state at its
// lower face, weno_left(v4..v0).  The mirrored stencil's smoothness
// indicators are the same three in reverse order, so the indicators, tau5
// and the ratios are computed once; each side keeps its own candidate
// polynomials and linear weights.
__device__ __forceinline__ void weno_pair(float v0, float v1, float v2, float v3, float v4,
                                          float eps, float& up, float& lo)
{
    float a = fmaf(-2.f, v1, v0) + v2;
    float b = fmaf(3.f, v2, fmaf(-4.f, v1, v0));
    const float b0 = fmaf((13.f / 12.f) * a, a, 0.25f * (b * b));
    a = fmaf(-2.f, v2, v1) + v3;
    b = v1 - v3;
    const float b1 = fmaf((13.f / 12.f) * a, a, 0.25f * (b * b));
    a = fmaf(-2.f, v3, v2) + v4;
    b = fmaf(-4.f, v3, fmaf(3.f, v2, v4));
    const float b2 = fmaf((13.f / 12.f) * a, a, 0.25f * (b * b));
    const float tau = fabsf(b2 - b0);
    const float r0 = fminf(tau * rcpf(b0 + eps), 1e18f);
    const float r1 = fminf(tau * rcpf(b1 + eps), 1e18f);
    const float r2 = fminf(tau * rcpf(b2 + eps), 1e18f);
    const float a1 = fmaf(0.6f * r1, r1, 0.6f);
    {
        const float p0 = fmaf(v0, 1.f / 3.f, fmaf(v1, -7.f / 6.f, v2 * (11.f / 6.f)));
        const float p1 = fmaf(v2, 5.f / 6.f, fmaf(v1, -1.f / 6.f, v3 * (1.f / 3.f)));
        const float p2 = fmaf(v2, 1.f / 3.f, fmaf(v3, 5.f / 6.f, v4 * (-1.f / 6.f)));
        const float a0 = fmaf(0.1f * r0, r0, 0.1f);
        const float a2 = fmaf(0.3f * r2, r2, 0.3f);
        up = fmaf(a0, p0, fmaf(a1, p1, a2 * p2)) * rcpf(a0 + a1 + a2);
    }
    {
        const float p0 = fmaf(v4, 1.f / 3.f, fmaf(v3, -7.f / 6.f, v2 * (11.f / 6.f)));
        const float p1 = fmaf(v2, 5.f / 6.f, fmaf(v3, -1.f / 6.f, v1 * (1.f / 3.f)));
        const float p2 = fmaf(v2, 1.f / 3.f, fmaf(v1, 5.f / 6.f, v0 * (-1.f / 6.f)));
        const float a0 = fmaf(0.1f * r2, r2, 0.1f);
        const float a2 = fmaf(0.3f * r0, r0, 0.3f);
        lo = fmaf(a0, p0, fmaf(a1, p1, a2 * p2)) * rcpf(a0 + a1 + a2);
    }
}

// Realizability projection in double (realizability.py:132-176).
__device__ __forceinline__ void project(double& p, double& x, double& y, double& z, double& jn)
{
    if (p < 0.0)
        p = 0.0;
    jn = sqrt(x * x + y * y + z * z);
    const double jmax = (1.0 - REAL_EPS) * p;
    if (jn > jmax) {
        const double s = jmax / jn;
        x *= s;
        y *= s;
        z *= s;
        jn = jmax;
    }
}

__device__ __forceinline__ void project_f(float u[4])
{
    double p = u[0], x = u[1], y = u[2], z = u[3], jn;
    project(p, x, y, z, jn);
    u[0] = float(p);
    u[1] = float(x);
    u[2] = float(y);
    u[3] = float(z);
}

// ---------------------------------------------------------------- float faces
// The float face path keeps every operation in float except the one quantity
// that loses its digits as |J|/phi -> 1: m = 1 - |J|^2/phi^2.  It enters the
// closure through 1 - chi = 2 m / (1 + sqrt(1 + 3 m)) (the Levermore closure
// with 4 - 3 f^2 = 1 + 3 m) and the eigenvalue discriminant, which the
// identity 1 - f^2 = a + 3 a^2 / 4 (a = 1 - chi) turns into the sum of
// non-negative terms 4 a (1 - mu^2) + 6 a^2 (1 + mu^2).  m itself is formed
// from error-free products and sums (fma and TwoSum), so it is exact to a few
// float ulps of m even where phi^2 - |J|^2 cancels.

// 1 - (1 - 1e-10)^2: states beyond it are on the realizability margin of
// realizability.py:182 (the reference scales them to |J| = (1 - 1e-10) phi).
constexpr float M_MIN = 2.0e-10f;

// phi^2 - |J|^2 for phi, J already scaled to phi in [0.5, 1).  The error
// terms are exact because no product here is fused into a later sum
// (-ffp-contract=off).
__device__ __forceinline__ float deficit(float p, float x, float y, float z)
{
    const float p2 = p * p, p2e = __builtin_fmaf(p, p, -p2);
    const float x2 = x * x, x2e = __builtin_fmaf(x, x, -x2);
    const float y2 = y * y, y2e = __builtin_fmaf(y, y, -y2);
    const float z2 = z * z, z2e = __builtin_fmaf(z, z, -z2);
    const float t = x2 + y2;
    float bp = t - x2;
    const float te = (x2 - (t - bp)) + (y2 - bp);
    const float u = t + z2;
    bp = u - t;
    const float ue = (t - (u - bp)) + (z2 - bp);
    return (p2 - u) + ((((p2e - x2e) - y2e) - z2e) - (te + ue));
}

// Projected face state with its closure quantities (realizability.py:132-176,
// flux.py:10-61): phi >= 0, |J| <= phi, m = 1 - f^2 >= M_MIN, a = 1 - chi.
// No reciprocal of phi or |J| is formed, so tiny and subnormal states stay
// finite.
struct FaceF {
    float p, J[3], n[3], jn, f, a;  // n = J / |J| (0 for J = 0)
};

__device__ __forceinline__ FaceF face_state(float q0, float q1, float q2, float q3)
{
    FaceF c;
    const float p = fmaxf(q0, 0.f);
    c.p = p;
    if (!(p > 0.f)) {
        // vacuum: J is projected to 0, f = 0, chi = 1/3
        c.J[0] = c.J[1] = c.J[2] = 0.f;
        c.n[0] = c.n[1] = c.n[2] = 0.f;
        c.jn = c.f = 0.f;
        c.a = 2.f / 3.f;
        return c;
    }
    // Exact power-of-two scaling by the largest magnitude keeps every square
    // below 1 (and phi in [0.5, 1) whenever |J| < phi).
    const float big = fmaxf(fmaxf(p, fabsf(q1)), fmaxf(fabsf(q2), fabsf(q3)));
    const int e = __builtin_amdgcn_frexp_expf(big);
    const float ps = __builtin_ldexpf(p, -e), xs = __builtin_ldexpf(q1, -e),
                ys = __builtin_ldexpf(q2, -e), zs = __builtin_ldexpf(q3, -e);
    const float d = deficit(ps, xs, ys, zs);
    // Direction from the dominant component k: with t = J / J_k,
    // n = t / |t| and |t|^2 = 1 + (the two ratios squared).  Where the other
    // two components are below ~2.4e-4 of J_k, |t|^2 rounds to 1 and n_k is
    // +-1 exactly, so a beam with J = phi along an axis keeps J = phi bit for
    // bit through fluxes and projections.
    const float ax = fabsf(xs), ay = fabsf(ys), az = fabsf(zs);
    const bool kz = az >= ax && az >= ay;
    const bool ky = !kz && ay >= ax;
    const float jk = kz ? zs : (ky ? ys : xs);
    const float u1 = kz ? xs : (ky ? xs : ys);  // the two other components,
    const float u2 = kz ? ys : (ky ? zs : zs);  // in axis order
    const bool nz = jk != 0.f;
    // |u| <= |J_k|: a second exact scaling by J_k's exponent keeps the
    // reciprocal finite however small J is
    const int ek = nz ? __builtin_amdgcn_frexp_expf(jk) : 0;
    const float ik = nz ? rcpf(__builtin_ldexpf(jk, -ek)) : 0.f;
    const float t1 = __builtin_ldexpf(u1, -ek) * ik, t2 = __builtin_ldexpf(u2, -ek) * ik;
    const float tt = 1.f + (t1 * t1 + t2 * t2);
    const float r = tt == 1.f ? 1.f : __builtin_amdgcn_rsqf(tt);
    const float nk = (jk > 0.f ? 1.f : -1.f) * r;
    const float n1 = t1 * nk, n2 = t2 * nk;  // t carries the sign of J_k
    c.n[0] = kz ? n1 : (ky ? n1 : nk);
    c.n[1] = kz ? n2 : (ky ? nk : n1);
    c.n[2] = kz ? nk : n2;
    if (!nz)
        c.n[0] = c.n[1] = c.n[2] = 0.f;
    const float jns = nz ? fabsf(jk) * (tt == 1.f ? 1.f : tt * r) : 0.f;
    if (d < M_MIN * (ps * ps)) {
        // on or beyond the margin: |J| scaled to phi, f = 1 - 1e-10 to float
        c.J[0] = p * c.n[0];
        c.J[1] = p * c.n[1];
        c.J[2] = p * c.n[2];
        c.jn = p;
        c.f = 1.f;
        c.a = 2.f * M_MIN * rcpf(1.f + __builtin_amdgcn_sqrtf(1.f + 3.f * M_MIN));
        return c;
    }
    // here |J| < phi, so big = phi and ps lies in [0.5, 1)
    const float inv_ps = rcpf(ps);
    const float m = d * (inv_ps * inv_ps);
    c.J[0] = q1;
    c.J[1] = q2;
    c.J[2] = q3;
    c.jn = __builtin_ldexpf(jns, e);
    c.f = fminf(jns * inv_ps, 1.f);
    c.a = 2.f * m * rcpf(1.f + __builtin_amdgcn_sqrtf(1.f + 3.f * m));
    return c;
}

// Stage-state projection in float (realizability.py:132-176).
__device__ __forceinline__ void project_ff(float u[4])
{
    const FaceF c = face_state(u[0], u[1], u[2], u[3]);
    u[0] = c.p;
    u[1] = c.J[0];
    u[2] = c.J[1];
    u[3] = c.J[2];
}

// Extreme eigenvalues along AXIS (wave_speeds.py:37-103) from the closure:
// disc = 4 a (1 - mu^2) + 6 a^2 (1 + mu^2), lambda = (2 f mu -+ sqrt(disc)) / (2 + 3 a).
template <int AXIS>
__device__ __forceinline__ void eigen_f(const FaceF& c, float eps, float& lmin, float& lmax)
{
    if (c.jn < eps) {
        lmin = -0.57735026918962576f;
        lmax = 0.57735026918962576f;
        return;
    }
    const float n0 = c.n[AXIS];
    const float n1 = c.n[(AXIS + 1) % 3];
    const float n2 = c.n[(AXIS + 2) % 3];
    const float mu2 = n0 * n0;
    const float disc = 4.f * c.a * (n1 * n1 + n2 * n2) + 6.f * c.a * c.a * (1.f + mu2);
    const float root = __builtin_amdgcn_sqrtf(disc);
    const float inv = rcpf(2.f + 3.f * c.a);
    lmin = (2.f * c.f * n0 - root) * inv;
    lmax = (2.f * c.f * n0 + root) * inv;
}

template <int AXIS>
__device__ __forceinline__ void phys_flux_f(const FaceF& c, float F[4])
{
    const float iso = 0.5f * c.a, dir = 1.f - 1.5f * c.a;
    const float na = c.n[AXIS];
    F[0] = c.J[AXIS];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        F[1 + k] = c.p * (dir * na * c.n[k] + (k == AXIS ? iso : 0.f));
}

struct Closure {
    double jn, f, chi, n[3], iso, dir;
};

// Closure of a projected state with |J| known (flux.py:10-61).
__device__ __forceinline__ Closure closure(double phi, const double J[3], double jn)
{
    Closure c;
    c.jn = jn;
    double f = phi > 0.0 ? jn / phi : 0.0;
    f = fmin(fmax(f, 0.0), 1.0);
    c.f = f;
    c.chi = (5.0 - 2.0 * sqrt(4.0 - 3.0 * f * f)) * (1.0 / 3.0);
    const double inv = jn > 0.0 ? 1.0 / jn : 0.0;
    c.n[0] = J[0] * inv;
    c.n[1] = J[1] * inv;
    c.n[2] = J[2] * inv;
    c.iso = 0.5 * (1.0 - c.chi);
    c.dir = 0.5 * (3.0 * c.chi - 1.0);
    return c;
}

// Extreme eigenvalues along AXIS (wave_speeds.py:37-103) with the scaled floor.
template <int AXIS>
__device__ __forceinline__ void eigen(double phi, const Closure& c, double eps, double& lmin,
                                      double& lmax)
{
    const double iso = 0.57735026918962576;
    if (c.jn < eps) {
        lmin = -iso;
        lmax = iso;
        return;
    }
    double f, chi;
    if (phi >= eps && c.f <= 1.0 - 1.0e-10) {
        f = c.f;
        chi = c.chi;
    } else {
        f = fmin(c.jn / fmax(phi, eps), 1.0 - 1.0e-10);
        chi = (5.0 - 2.0 * sqrt(4.0 - 3.0 * f * f)) * (1.0 / 3.0);
    }
    const double mu = c.n[AXIS];
    const double den = 5.0 - 3.0 * chi;
    const double disc =
        fmax(4.0 * f * f * mu * mu - den * (mu * mu * (3.0 * chi - 1.0) - 2.0 * (1.0 - chi)), 0.0);
    const double root = sqrt(disc);
    const double inv = 1.0 / den;
    lmin = (2.0 * f * mu - root) * inv;
    lmax = (2.0 * f * mu + root) * inv;
}

template <int AXIS>
__device__ __forceinline__ void phys_flux(double phi, const double J[3], const Closure& c,
                                          double F[4])
{
    F[0] = J[AXIS];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        F[1 + k] = phi * (c.dir * c.n[AXIS] * c.n[k] + (k == AXIS ? c.iso : 0.0));
}

// HLL flux of two projected float face states (flux.py:358-384).
template <int AXIS>
__device__ __forceinline__ void hll_f(const FaceF& cl, const FaceF& cr, int wave, float eps,
                                      float F[4])
{
    float FL[4], FR[4];
    phys_flux_f<AXIS>(cl, FL);
    phys_flux_f<AXIS>(cr, FR);
    float sl = -1.f, sr = 1.f;
    if (wave == 0) {
        float a, b, c, d;
        eigen_f<AXIS>(cl, eps, a, b);
        eigen_f<AXIS>(cr, eps, c, d);
        sl = fminf(fminf(a, c), 0.f);
        sr = fmaxf(fmaxf(b, d), 0.f);
    }
    const float inv = rcpf(sr - sl);
    const float sp = sl * sr;
    const float Lp[4] = {cl.p, cl.J[0], cl.J[1], cl.J[2]};
    const float Rp[4] = {cr.p, cr.J[0], cr.J[1], cr.J[2]};
#pragma unroll
    for (int c = 0; c < 4; ++c)
        F[c] = (sr * FL[c] - sl * FR[c] + sp * (Rp[c] - Lp[c])) * inv;
}

// HLL flux of a face from its reconstructed left and right states.
// D: projection, closure, eigenvalues and HLL in double; otherwise in float
// with the cancellation-free closure above.
template <int AXIS, bool D>
__device__ __forceinline__ void face_hll(const float Ls[4], const float Rs[4], int wave,
                                         double eeps, float F[4])
{
    if constexpr (D) {
        double L[4], R[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            L[c] = Ls[c];
            R[c] = Rs[c];
        }
        double jl, jr;
        project(L[0], L[1], L[2], L[3], jl);
        project(R[0], R[1], R[2], R[3], jr);
        const Closure cl = closure(L[0], L + 1, jl);
        const Closure cr = closure(R[0], R + 1, jr);
        double FL[4], FR[4];
        phys_flux<AXIS>(L[0], L + 1, cl, FL);
        phys_flux<AXIS>(R[0], R + 1, cr, FR);
        double sl = -1.0, sr = 1.0;
        if (wave == 0) {
            double a, b, c, d;
            eigen<AXIS>(L[0], cl, eeps, a, b);
            eigen<AXIS>(R[0], cr, eeps, c, d);
            sl = fmin(fmin(a, c), 0.0);
            sr = fmax(fmax(b, d), 0.0);
        }
        const double inv = 1.0 / (sr - sl);
        const double sp = sl * sr;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            F[c] = float((sr * FL[c] - sl * FR[c] + sp * (R[c] - L[c])) * inv);
    } else {
        const FaceF cl = face_state(Ls[0], Ls[1], Ls[2], Ls[3]);
        const FaceF cr = face_state(Rs[0], Rs[1], Rs[2], Rs[3]);
        hll_f<AXIS>(cl, cr, wave, float(eeps), F);
    }
}

// Stencil input of one cell: U directly, or proj(C_k V_k / S_k) for stages 1, 2.
// PRE: stages 1 and 2 read the input the previous stage formed (a.win).
template <int KIND, bool D, bool PRE>
__device__ __forceinline__ void load_state(const TileView& a, const StepCoef32& c, long i,
                                           long N, float u[4])
{
    if constexpr (PRE && (KIND == STAGE1 || KIND == STAGE2)) {
        u[0] = a.win[i];
        u[1] = a.win[i + N];
        u[2] = a.win[i + 2 * N];
        u[3] = a.win[i + 3 * N];
        return;
    }
    const float* src = a.src;
    if constexpr (KIND == STAGE_RELAX) {
        // quasi-steady cells (a nonzero pseudo-time CFL number) from the
        // previous sweep's output, when there is one
        if (a.V1 && c.dEm[a.ids[i]] != 0.f)
            src = a.V1;
    }
    u[0] = src[i];
    u[1] = src[i + N];
    u[2] = src[i + 2 * N];
    u[3] = src[i + 3 * N];
    if constexpr (KIND == STAGE1 || KIND == STAGE2) {
        const int m = a.ids[i];
        const float gp = KIND == STAGE1 ? c.g1p[m] : c.g2p[m];
        const float gj = KIND == STAGE1 ? c.g1j[m] : c.g2j[m];
        u[0] *= gp;
        u[1] *= gj;
        u[2] *= gj;
        u[3] *= gj;
        if constexpr (D)
            project_f(u);
        else
            project_ff(u);
    }
}

// Quartic z-min ghost k (boundary.py:274-299) in double.
__device__ __forceinline__ float zmin_ghost(int k, double uin, double u0, double u1, double u2,
                                            double u3)
{
    double g;
    if (k == 1)
        g = 5.0 * uin - (77.0 / 12.0) * u0 + (43.0 / 12.0) * u1 - (17.0 / 12.0) * u2 + 0.25 * u3;
    else if (k == 2)
        g = 25.0 * uin - (505.0 / 12.0) * u0 + (335.0 / 12.0) * u1 - (145.0 / 12.0) * u2 +
            (9.0 / 4.0) * u3;
    else
        g = 75.0 * uin - (545.0 / 4.0) * u0 + (395.0 / 4.0) * u1 - (181.0 / 4.0) * u2 +
            (35.0 / 4.0) * u3;
    return float(g);
}

__device__ __forceinline__ double amax(double lo, double hi) { return fmax(fabs(lo), fabs(hi)); }

// Inverse CFL step of one cell of the scaled float state (energy.py:539-660):
// sum over axes of max|lambda| / (S d_axis).  The speeds of the cell's state
// come from the float closure of the stage kernels (face_state, eigen_f: the
// state is already projected, so face_state only forms the closure); a z-min
// cell also counts the inflow state (p_in = scaled inflow phi,
// J = (0, 0, rf p_in)) on the z axis, in double.
__device__ __forceinline__ double cfl_cell(const float* U, long i, long N, double s,
                                           const GridDims& g, int wave_mode, double eps,
                                           bool zmin, double p_in, double rf)
{
    if (wave_mode == 1)
        return 1.0 / (s * g.dx) + 1.0 / (s * g.dy) + 1.0 / (s * g.dz);
    const FaceF f = face_state(U[i], U[i + N], U[i + 2 * N], U[i + 3 * N]);
    const float epsf = float(eps);
    float l, h;
    eigen_f<0>(f, epsf, l, h);
    const double ax = fmaxf(fabsf(l), fabsf(h));
    eigen_f<1>(f, epsf, l, h);
    const double ay = fmaxf(fabsf(l), fabsf(h));
    eigen_f<2>(f, epsf, l, h);
    double az = fmaxf(fabsf(l), fabsf(h));
    double lo, hi;
    if (zmin) {
        const double Jin[3] = {0.0, 0.0, rf * p_in};
        const Closure ci = closure(p_in, Jin, fabs(Jin[2]));
        eigen<2>(p_in, ci, eps, lo, hi);
        az = fmax(az, amax(lo, hi));
    }
    return ax / (s * g.dx) + ay / (s * g.dy) + az / (s * g.dz);
}

// cfl_cell with the directional masks of the vacuum-aware CFL
// (energy.py:610-636): an axis whose mask is off contributes nothing; the
// z-min inflow speed joins the z axis before its mask is applied.
__device__ __forceinline__ double cfl_cell_masked(const float* U, long i, long N, double s,
                                                  const GridDims& g, double eps, bool zmin,
                                                  double p_in, double rf, bool mx, bool my,
                                                  bool mz)
{
    if (!(mx || my || mz))
        return 0.0;
    const double phi = U[i];
    const double J[3] = {U[i + N], U[i + 2 * N], U[i + 3 * N]};
    const double jn = sqrt(J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
    Closure c = closure(phi, J, jn);
    double lo, hi, ax = 0.0, ay = 0.0, az = 0.0;
    if (mx) {
        eigen<0>(phi, c, eps, lo, hi);
        ax = amax(lo, hi);
    }
    if (my) {
        eigen<1>(phi, c, eps, lo, hi);
        ay = amax(lo, hi);
    }
    if (mz) {
        eigen<2>(phi, c, eps, lo, hi);
        az = amax(lo, hi);
        if (zmin) {
            const double Jin[3] = {0.0, 0.0, rf * p_in};
            const Closure ci = closure(p_in, Jin, fabs(Jin[2]));
            eigen<2>(p_in, ci, eps, lo, hi);
            az = fmax(az, amax(lo, hi));
        }
    }
    return ax / (s * g.dx) + ay / (s * g.dy) + az / (s * g.dz);
}

// Inverse CFL step of a cell holding the exact zero state: the isotropic
// speeds 1/sqrt(3) on every axis (wave_speeds.py:37-50), the same value
// cfl_cell returns for U = 0.
__device__ __forceinline__ double cfl_vacuum(double s, const GridDims& g, int wave_mode, double eps,
                                             bool zmin, double p_in, double rf)
{
    const double iso = 0.57735026918962576f;  // eigen_f's isotropic speed, in float
    if (wave_mode == 1)
        return 1.0 / (s * g.dx) + 1.0 / (s * g.dy) + 1.0 / (s * g.dz);
    double az = iso;
    if (zmin) {
        double lo, hi;
        const double Jin[3] = {0.0, 0.0, rf * p_in};
        const Closure ci = closure(p_in, Jin, fabs(Jin[2]));
        eigen<2>(p_in, ci, eps, lo, hi);
        az = fmax(az, amax(lo, hi));
    }
    return iso / (s * g.dx) + iso / (s * g.dy) + az / (s * g.dz);
}

// Workgroup maximum of mx (NaN if any thread saw one) folded into *out as
// ordered bits (non-negative doubles order like their bit patterns; NaN
// sorts above every finite value).  Needs blockDim.x / 64 slots of `red`.
__device__ __forceinline__ void block_max_to(double mx, bool bad, double* red,
                                             unsigned long long* out)
{
    if (bad)
        mx = __longlong_as_double(0x7ff8000000000000LL);
    unsigned long long bits = __double_as_longlong(mx);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = __longlong_as_double(bits);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = 0;
        for (int w = 0; w < int(blockDim.x / 64); ++w) {
            const unsigned long long o = __double_as_longlong(red[w]);
            b = o > b ? o : b;
        }
        atomicMax(out, b);
    }
}

// Stage boundary traces of one z-min column b (energy.py:107-230, 348-455) in
// double, stored scaled by 1/alpha in float: bnd holds 3 x (4, m2) values for
// stages 0, 1, 2.  m is the material of the column's z = 0 cell and sp its
// inflow profile value.
__device__ __forceinline__ void boundary_column(const StepCoef& c, int m, double sp, long b,
                                                long m2, bool integrating_factor, float* bnd)
{
    const double ia = c.inv_alpha;
    float* o0 = bnd;
    float* o1 = bnd + 4 * m2;
    float* o2 = bnd + 8 * m2;
    const double p0 = sp * c.ed[0];
    o0[b] = float(p0 * ia);
    o0[m2 + b] = 0.f;
    o0[2 * m2 + b] = 0.f;
    o0[3 * m2 + b] = float(c.rf * p0 * ia);
    const double Sk[4] = {c.S0[m], c.S13[m], c.S23[m], c.S1[m]};
    double B[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double p = sp * c.ed[k];
        B[k][0] = Sk[k] * p;
        B[k][1] = 0.0;
        B[k][2] = 0.0;
        B[k][3] = Sk[k] * (c.rf * p);
    }
    if (integrating_factor) {
        const double ip[3] = {c.i13p[m], c.i23p[m], c.i1p[m]};
        const double ij[3] = {c.i13j[m], c.i23j[m], c.i1j[m]};
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            B[k][0] *= ip[k - 1];
            B[k][3] *= ij[k - 1];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v1 = 0.5 * (-9.0 * B[0][q] + 18.0 * B[1][q] - 9.0 * B[2][q] + 2.0 * B[3][q]);
        double v2 = 0.25 * (11.0 * B[0][q] - 27.0 * B[1][q] + 27.0 * B[2][q] - 7.0 * B[3][q]);
        if (integrating_factor) {
            v1 *= (q == 0 ? c.f1p[m] : c.f1j[m]);
            v2 *= (q == 0 ? c.f2p[m] : c.f2j[m]);
        }
        o1[q * m2 + b] = float(v1 / c.S1[m] * ia);
        o2[q * m2 + b] = float(v2 / c.S2[m] * ia);
    }
}

// Compensated accumulation hi + lo += v (Knuth TwoSum).
__device__ __forceinline__ void two_sum_add(float& hi, float& lo, float v)
{
    const float s = hi + v;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (v - bp);
    hi = s;
    lo += err;
}

// The next stage's stencil input of one cell from this stage's stored
// output v: proj(g v), g = C_k / S_k of the cell's material -- load_state's
// arithmetic on the same floats, so the next stage reads what it would form.
template <bool D>
__device__ __forceinline__ void store_next_input(float* w, long i, long N, const float v[4],
                                                 float gp, float gj)
{
    float u[4] = {v[0] * gp, v[1] * gj, v[2] * gj, v[3] * gj};
    if constexpr (D)
        project_f(u);
    else
        project_ff(u);
#pragma unroll
    for (int c = 0; c < 4; ++c)
        w[i + c * N] = u[c];
}

// PRE: stages 0 and 1 also store the next stage's input (a.wout) and stages
// 1 and 2 read theirs (a.win) instead of forming it from src.
template <int KIND, bool IFAC, bool D, bool PRE = false>
__device__ __forceinline__ void stage_tile(const TileView& a)
{
    __shared__ StepCoef32 sc;
    __shared__ float P[4][PY][PZ];
    // Fy / Fz hold the upper WENO states of the faces, then their fluxes;
    // Hy / Hz the lower states of the cells just above the tile.
    __shared__ float Fy[4][TY + 1][TZ];
    __shared__ float Fz[4][TY][TZ + 1];
    __shared__ float Hy[4][TZ];
    __shared__ float Hz[4][TY];
    // per thread: lower states of its cell along z and y (until the fluxes)
    // and the upper x state carried to the next plane
    __shared__ float Rzs[4][TY][TZ], Rys[4][TY][TZ], Lxs[4][TY][TZ];
    __shared__ float wmax[3][NT / 64];
    __shared__ float plane[NT / 64][TX];  // per wave, max phi of each x plane
    __shared__ float colmax[TY][TZ];      // per (y, z) column, max phi over x

    const int tid = threadIdx.x;
    float phimax = 0.f, phisum = 0.f, jsum = 0.f;
    {
        const unsigned* src = reinterpret_cast<const unsigned*>(a.coef);
        unsigned* dst = reinterpret_cast<unsigned*>(&sc);
        for (int k = tid; k < int(sizeof(StepCoef32) / 4); k += NT)
            dst[k] = src[k];
    }
    __syncthreads();

    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long N = long(nx) * ny * nz;
    const int tz = tid % TZ, ty = tid / TZ;
    const int z0 = a.tz * TZ, y0 = a.ty * TY, x0 = a.tx * TX;
    const int z = z0 + tz, y = y0 + ty;
    const int zc = min(z, nz - 1), yc = min(y, ny - 1);
    const bool valid = z < nz && y < ny;
    const int x1 = min(x0 + TX, nx);
    const int wave = a.wave_mode;
    const float weps = sc.weno_eps;
    const double eeps = sc.eig_eps;
    const float idx = float(1.0 / a.g.dx), idy = float(1.0 / a.g.dy), idz = float(1.0 / a.g.dz);

    auto cidx = [&](int xx, int yy, int zz) -> long { return (long(xx) * ny + yy) * nz + zz; };
    auto clampx = [&](int xx) { return xx < 0 ? 0 : (xx >= nx ? nx - 1 : xx); };

    int hy = -1, hz = -1;
    if (tid < 6 * TZ) {
        const int r = tid / TZ, col = tid % TZ;
        hy = r < 3 ? r : TY + r;
        hz = col + 3;
    } else if (tid < NHALO) {
        const int h = tid - 6 * TZ;
        const int r = h / 6, c6 = h % 6;
        hy = r + 3;
        hz = c6 < 3 ? c6 : TZ + c6;
    }
    const int hgy = min(max(y0 + hy - 3, 0), ny - 1);
    const int hgz_raw = z0 + hz - 3;
    const int hgz = min(max(hgz_raw, 0), nz - 1);

    // x: a window of six planes of this thread's column; each cell's two
    // WENO states are computed once, the left state of its upper face
    // carried to the next plane.
    float W[6][4];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        load_state<KIND, D, PRE>(a, sc, cidx(clampx(x0 - 3 + k), yc, zc), N, W[k]);
    float Flo[4];
    {
        float L[4], R[4], up;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            L[c] = weno_left(W[0][c], W[1][c], W[2][c], W[3][c], W[4][c], weps);
            weno_pair(W[1][c], W[2][c], W[3][c], W[4][c], W[5][c], weps, up, R[c]);
            Lxs[c][ty][tz] = up;
        }
        face_hll<0, D>(L, R, wave, eeps, Flo);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            W[k][c] = W[k + 1][c];

    // y and z: lanes of the halo cells' states next to the tile (upper state
    // of the cell below the tile, lower state of the cell above it).
    const int nh = 2 * (TY + TZ);
    const bool hz_lane = tid < 2 * TY;
    const int h_line = hz_lane ? tid % TY : (tid - 2 * TY) % TZ;
    const bool h_high = hz_lane ? tid >= TY : tid - 2 * TY >= TZ;

    for (int x = x0; x < x1; ++x) {
        load_state<KIND, D, PRE>(a, sc, cidx(clampx(x + 3), yc, zc), N, W[5]);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            P[c][ty + 3][tz + 3] = W[2][c];
        if (hy >= 0) {
            float u[4];
            if (hgz_raw < 0) {
                float q[4][4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    load_state<KIND, D, PRE>(a, sc, cidx(x, hgy, k), N, q[k]);
                const long b = long(x) * ny + hgy;
                const long m2 = long(nx) * ny;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    u[c] = zmin_ghost(-hgz_raw, a.bnd[c * m2 + b], q[0][c], q[1][c], q[2][c],
                                      q[3][c]);
            } else {
                load_state<KIND, D, PRE>(a, sc, cidx(x, hgy, hgz), N, u);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
                P[c][hy][hz] = u[c];
        }
        float Fhi[4];
        {
            float L[4], R[4], up;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                L[c] = Lxs[c][ty][tz];
                weno_pair(W[1][c], W[2][c], W[3][c], W[4][c], W[5][c], weps, up, R[c]);
                Lxs[c][ty][tz] = up;
            }
            face_hll<0, D>(L, R, wave, eeps, Fhi);
        }
        __syncthreads();
        // This cell's states along z and y; the upper states go to LDS (the
        // flux arrays, slot of the face they belong to) for the cell above.
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            float up, lo;
            weno_pair(P[c][ty + 3][tz + 1], P[c][ty + 3][tz + 2], P[c][ty + 3][tz + 3],
                      P[c][ty + 3][tz + 4], P[c][ty + 3][tz + 5], weps, up, lo);
            Fz[c][ty][tz + 1] = up;
            Rzs[c][ty][tz] = lo;
            weno_pair(P[c][ty + 1][tz + 3], P[c][ty + 2][tz + 3], P[c][ty + 3][tz + 3],
                      P[c][ty + 4][tz + 3], P[c][ty + 5][tz + 3], weps, up, lo);
            Fy[c][ty + 1][tz] = up;
            Rys[c][ty][tz] = lo;
        }
        if (tid < nh) {
            // z lanes: row h_line, centre z = -1 (upper state) or TZ (lower);
            // y lanes: column h_line, centre y = -1 or TY
#pragma unroll 1
            for (int c = 0; c < 4; ++c) {
                // the upper cell's lower state is the mirrored stencil's left
                // state: reverse the operands, so that every lane runs one
                // weno_left (the lanes of both kinds share a wave)
                float v[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const int kk = h_high ? 4 - k : k;
                    v[k] = hz_lane ? P[c][h_line + 3][(h_high ? TZ + 1 : 0) + kk]
                                   : P[c][(h_high ? TY + 1 : 0) + kk][h_line + 3];
                }
                const float st = weno_left(v[0], v[1], v[2], v[3], v[4], weps);
                if (hz_lane) {
                    if (h_high)
                        Hz[c][h_line] = st;
                    else
                        Fz[c][h_line][0] = st;
                } else {
                    if (h_high)
                        Hy[c][h_line] = st;
                    else
                        Fy[c][0][h_line] = st;
                }
            }
        }
        __syncthreads();
        // Fluxes: each cell's lower faces, then the upper faces of the tile.
        {
            float L[4], R[4], F[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                L[c] = Fz[c][ty][tz];
                R[c] = Rzs[c][ty][tz];
            }
            face_hll<2, D>(L, R, wave, eeps, F);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Fz[c][ty][tz] = F[c];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                L[c] = Fy[c][ty][tz];
                R[c] = Rys[c][ty][tz];
            }
            face_hll<1, D>(L, R, wave, eeps, F);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Fy[c][ty][tz] = F[c];
        }
        // Upper faces of the tile: the z faces in the first wave, the y
        // faces in the second, so that no wave runs both flux variants.
        if (tid < TY) {
            float L[4], R[4], F[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                L[c] = Fz[c][tid][TZ];
                R[c] = Hz[c][tid];
            }
            face_hll<2, D>(L, R, wave, eeps, F);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Fz[c][tid][TZ] = F[c];
        } else if (tid >= 64 && tid < 64 + TZ) {
            const int line = tid - 64;
            float L[4], R[4], F[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                L[c] = Fy[c][TY][line];
                R[c] = Hy[c][line];
            }
            face_hll<1, D>(L, R, wave, eeps, F);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Fy[c][TY][line] = F[c];
        }
        __syncthreads();

        float xphi = 0.f;  // STAGE2: phi of this cell, for the face slabs
        if (valid) {
            const long i = cidx(x, y, z);
            const int m = a.ids[i];
            const float dE = sc.dE;      // the step's energy width (dose)
            const float dEr = sc.dEm[m];  // the step on the divergence (0: held)
            float rhs[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                rhs[c] = -fmaf(Fhi[c] - Flo[c], idx,
                               fmaf(Fy[c][ty + 1][tz] - Fy[c][ty][tz], idy,
                                    (Fz[c][ty][tz + 1] - Fz[c][ty][tz]) * idz));
            if constexpr (KIND == STAGE_RHS) {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    a.out[i + c * N] = rhs[c];
            } else if constexpr (KIND == STAGE_RELAX) {
                // a pseudo-time step towards div F = 0 of a quasi-steady cell,
                // dtau = cfl / (sum over the axes of its speed / cell size)
                if (dEr != 0.f) {
                    const double sp = cfl_cell(a.V1 ? a.V1 : a.src, i, N, 1.0, a.g, wave, eeps,
                                               false, 0.0, 1.0);
                    const float dtau = sp > 0.0 ? float(double(dEr) / sp) : 0.f;
                    float u[4];
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        u[c] = fmaf(rhs[c], dtau, W[2][c]);
                    if constexpr (D)
                        project_f(u);
                    else
                        project_ff(u);
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        a.out[i + c * N] = u[c];
                        phimax = fmaxf(phimax, fabsf(u[c] - W[2][c]));  // the sweep's change
                    }
                }
            } else if constexpr (KIND == STAGE0) {
                const float S0 = sc.S0[m];
                float v[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    v[c] = fmaf(rhs[c], dEr, S0 * W[2][c]);
                    a.out[i + c * N] = v[c];
                }
                if constexpr (PRE)
                    store_next_input<D>(a.wout, i, N, v, sc.g1p[m], sc.g1j[m]);
            } else if constexpr (KIND == STAGE_FE) {
                const float S0 = sc.S0[m];
                float u[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    u[c] = fmaf(rhs[c], dEr, S0 * W[2][c]) * (1.f / sc.S1[m]);
                if constexpr (D)
                    project_f(u);
                else
                    project_ff(u);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    a.out[i + c * N] = u[c];
                const float e = dE * W[2][0] * S0;
                if (a.edep)
                    a.edep[i] = e;
                if (a.dose_hi)
                    two_sum_add(a.dose_hi[i], a.dose_lo[i], e * sc.dose_scale[m]);
            } else if constexpr (KIND == STAGE1) {
                const float S0 = sc.S0[m];
                float v[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float r = rhs[c];
                    if constexpr (IFAC)
                        r *= (c == 0 ? sc.i1p[m] : sc.i1j[m]);
                    v[c] = fmaf(0.75f, S0 * a.U[i + c * N], 0.25f * fmaf(r, dEr, a.V1[i + c * N]));
                    a.out[i + c * N] = v[c];
                }
                if constexpr (PRE)
                    store_next_input<D>(a.wout, i, N, v, sc.g2p[m], sc.g2j[m]);
            } else {  // STAGE2
                const float S0 = sc.S0[m], S2 = sc.S2[m];
                float un[4], u[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    un[c] = a.U[i + c * N];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float r = rhs[c];
                    if constexpr (IFAC)
                        r *= (c == 0 ? sc.i2p[m] : sc.i2j[m]);
                    const float v = fmaf(1.f / 3.f, S0 * un[c],
                                         (2.f / 3.f) * fmaf(r, dEr, a.src[i + c * N]));
                    u[c] = v * (c == 0 ? sc.g1p[m] : sc.g1j[m]);
                }
                if constexpr (D)
                    project_f(u);
                else
                    project_ff(u);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    a.U[i + c * N] = u[c];
                phimax = fmaxf(phimax, u[0]);
                xphi = u[0];
                if (a.tile_sums) {
                    phisum += u[0];
                    jsum += sqrtf(fmaf(u[1], u[1], fmaf(u[2], u[2], u[3] * u[3])));
                }
                const float q0 = un[0] * S0;
                const float q1 = fmaxf(a.V1[i] * sc.g1p[m], 0.f) * sc.S1[m];
                const float q2 = W[2][0] * S2;
                const float e = dE * fmaf(1.f / 6.f, q0 + q1, (2.f / 3.f) * q2);
                if (a.edep)
                    a.edep[i] = e;
                if (a.dose_hi)
                    two_sum_add(a.dose_hi[i], a.dose_lo[i], e * sc.dose_scale[m]);
            }
        }
        if constexpr (KIND == STAGE2) {
            if (a.tile_face) {
                for (int off = 32; off > 0; off >>= 1)
                    xphi = fmaxf(xphi, __shfl_xor(xphi, off, 64));
                if ((tid & 63) == 0)
                    plane[tid >> 6][x - x0] = xphi;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            Flo[c] = Fhi[c];
#pragma unroll
        for (int k = 0; k < 5; ++k)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                W[k][c] = W[k + 1][c];
    }
    if constexpr (KIND == STAGE_RELAX) {
        // largest change of a quasi-steady cell this sweep, folded into the
        // box's slot as ordered bits (non-negative floats order like them)
        if (a.tile_max) {
            for (int off = 32; off > 0; off >>= 1)
                phimax = fmaxf(phimax, __shfl_xor(phimax, off, 64));
            if ((tid & 63) == 0)
                atomicMax(reinterpret_cast<unsigned*>(a.tile_max), __float_as_uint(phimax));
        }
    }
    if constexpr (KIND == STAGE2) {
        if (a.tile_max) {
            if (a.tile_face)
                colmax[ty][tz] = phimax;
            for (int off = 32; off > 0; off >>= 1) {
                phimax = fmaxf(phimax, __shfl_xor(phimax, off, 64));
                phisum += __shfl_xor(phisum, off, 64);
                jsum += __shfl_xor(jsum, off, 64);
            }
            if ((tid & 63) == 0) {
                wmax[0][tid >> 6] = phimax;
                wmax[1][tid >> 6] = phisum;
                wmax[2][tid >> 6] = jsum;
            }
            __syncthreads();
            if (tid == 0) {
                float m = wmax[0][0], ps = wmax[1][0], js = wmax[2][0];
                for (int w = 1; w < NT / 64; ++w) {
                    m = fmaxf(m, wmax[0][w]);
                    ps += wmax[1][w];
                    js += wmax[2][w];
                }
                *a.tile_max = m;
                if (a.tile_sums) {
                    a.tile_sums[0] = ps;
                    a.tile_sums[1] = js;
                }
            }
            if (a.tile_face && tid < FACE_SLOTS) {
                // slab tid: depth TILE_REACH (slots 0-5) or TILE_GUARD (6-11),
                // axis (tid % 6) / 2, low (even) or high (odd) side
                const int depth = tid < 6 ? TILE_REACH : TILE_GUARD;
                const int axis = (tid % 6) / 2;
                const bool high = tid & 1;
                const int ext = axis == 0 ? TX : (axis == 1 ? TY : TZ);
                const int lo = high ? max(ext - depth, 0) : 0;
                const int hi = high ? ext : min(depth, ext);
                float v = 0.f;
                if (axis == 0) {
                    const int nxl = x1 - x0;
                    for (int w = 0; w < NT / 64; ++w)
                        for (int q = lo; q < min(hi, nxl); ++q)
                            v = fmaxf(v, plane[w][q]);
                } else if (axis == 1) {
                    for (int q = lo; q < hi; ++q)
                        for (int r = 0; r < TZ; ++r)
                            v = fmaxf(v, colmax[q][r]);
                } else {
                    for (int q = 0; q < TY; ++q)
                        for (int r = lo; r < hi; ++r)
                            v = fmaxf(v, colmax[q][r]);
                }
                a.tile_face[tid] = v;
            }
        }
    }
}


}  // namespace detail
}  // namespace f32
}  // namespace m1
