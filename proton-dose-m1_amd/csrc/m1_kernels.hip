// Fused HIP kernels of one M1 energy step for gfx950 (FP64).
//
// One launch per SSPRK3 stage computes, for every cell of a tile, the
// WENO-Z reconstruction of all three axes, the per-face realizability
// projection, the HLL flux with the exact M1 wave speeds, the flux divergence
// and the stage update -- the whole of spatial_operator (spatial.py:104-252)
// plus the stage algebra of ssprk3 / if_ssprk3 (energy.py:725-1130) -- without
// materialising ghost arrays, face states or fluxes in HBM.
//
// Tiling: a workgroup of 256 threads owns a TY x TZ = 8 x 32 tile of a y-z
// plane and marches along x over TX planes.  The x-stencil of each thread's
// cell column lives in registers (a window of six planes); the y and z
// stencils are read from an LDS copy of the current plane with a 3-cell halo
// (the WENO-Z ghost width, weno_z.py:6).  Each face flux is computed once:
// every thread computes the lower y and z faces of its cell, 40 threads the
// upper faces of the tile, and the x face ahead of the march.
//
// Only the unprojected stage variables V1, V2 are stored; the projected stage
// state U_k = proj(C_k V_k / S_k) (energy.py:1005-1016, 1059-1070) is
// recomputed pointwise whenever a cell is loaded.
//
// Built twice, see m1_device.h: M1_EXACT=1 keeps the reference's rounding
// sequence, M1_EXACT=0 lets the compiler fuse multiply-adds and replaces
// divisions by constants with reciprocal multiplications.
#include "m1_device.h"

#ifndef M1_EXACT
#define M1_EXACT 0
#endif

#if M1_EXACT
#define M1_NS exact
#else
#define M1_NS fast
#endif

namespace m1 {
namespace M1_NS {
namespace {

constexpr int TY = 8;
constexpr int TZ = 32;
constexpr int TX = 16;
constexpr int NT = TY * TZ;
constexpr int PY = TY + 6;
constexpr int PZ = TZ + 6;
constexpr int NHALO = 6 * TZ + 6 * TY;
static_assert(NHALO <= NT, "one halo cell per thread");
static_assert(TY + TZ <= 64, "upper tile faces fit one wavefront");

constexpr double EIG_EPS = 1.0e-10;   // wave_speeds.py:10
constexpr double WENO_EPS = 1.0e-12;  // weno_z.py:8
constexpr double REAL_EPS = 1.0e-10;  // weno_z.py:10, realizability.py:182

// ------------------------------------------------------- exact helpers
// x / D for the constant divisors of the reconstruction and closure.  The
// quotient x * RN(1/D) corrected once with its exact FMA residual is the
// correctly rounded x / D for every normal x (Markstein's division by a
// constant with a correctly rounded reciprocal; checked on 8e8 random
// operands for D = 3, 4, 6, 12).  The residual can underflow for tiny x, and
// the sign of a zero quotient must follow x, so |x| < 2^-960, zeros and NaNs
// take the IEEE division.  The fast build multiplies by the reciprocal.
template <int D>
__device__ __forceinline__ double div_by(double x)
{
    constexpr double d = D;
    constexpr double r = 1.0 / D;
#if M1_EXACT
    double q = x * r;
    const double t = __builtin_fma(-q, d, x);
    q = __builtin_fma(t, r, q);
    if (__builtin_expect(!(fabs(x) >= 0x1p-960), 0))
        q = x / d;
    return q;
#else
    return x * r;
#endif
}

// ------------------------------------------------------------------ WENO-Z
// weno_z.py:13-87: left state at the face between v2 and v3.
__device__ __forceinline__ double weno_left(double v0, double v1, double v2, double v3,
                                            double v4)
{
#if M1_EXACT
    const double p0 = div_by<3>(v0) - div_by<6>(7.0 * v1) + div_by<6>(11.0 * v2);
    const double p1 = div_by<6>(-v1) + div_by<6>(5.0 * v2) + div_by<3>(v3);
    const double p2 = div_by<3>(v2) + div_by<6>(5.0 * v3) - div_by<6>(v4);
#else
    const double p0 = v0 * (1.0 / 3.0) - v1 * (7.0 / 6.0) + v2 * (11.0 / 6.0);
    const double p1 = v2 * (5.0 / 6.0) - v1 * (1.0 / 6.0) + v3 * (1.0 / 3.0);
    const double p2 = v2 * (1.0 / 3.0) + v3 * (5.0 / 6.0) - v4 * (1.0 / 6.0);
#endif
    double a = v0 - 2.0 * v1 + v2;
    double b = v0 - 4.0 * v1 + 3.0 * v2;
    const double beta0 = 13.0 / 12.0 * (a * a) + 0.25 * (b * b);
    a = v1 - 2.0 * v2 + v3;
    b = v1 - v3;
    const double beta1 = 13.0 / 12.0 * (a * a) + 0.25 * (b * b);
    a = v2 - 2.0 * v3 + v4;
    b = 3.0 * v2 - 4.0 * v3 + v4;
    const double beta2 = 13.0 / 12.0 * (a * a) + 0.25 * (b * b);
    const double tau5 = fabs(beta2 - beta0);
    const double r0 = tau5 / (beta0 + WENO_EPS);
    const double r1 = tau5 / (beta1 + WENO_EPS);
    const double r2 = tau5 / (beta2 + WENO_EPS);
    const double alpha0 = 0.1 * (1.0 + r0 * r0);
    const double alpha1 = 0.6 * (1.0 + r1 * r1);
    const double alpha2 = 0.3 * (1.0 + r2 * r2);
    const double s = alpha0 + alpha1 + alpha2;
#if M1_EXACT
    const double w0 = alpha0 / s, w1 = alpha1 / s, w2 = alpha2 / s;
    return w0 * p0 + w1 * p1 + w2 * p2;
#else
    const double inv = 1.0 / s;
    return (alpha0 * p0 + alpha1 * p1 + alpha2 * p2) * inv;
#endif
}

// weno_z.py:91-129 and realizability.py:132-176.
__device__ __forceinline__ void project(double& p, double& x, double& y, double& z)
{
    if (p < 0.0)
        p = 0.0;
    const double jn = sqrt(x * x + y * y + z * z);
    const double jmax = (1.0 - REAL_EPS) * p;
    if (jn > jmax) {
        const double s = jmax / jn;
        x *= s;
        y *= s;
        z *= s;
    }
}

// ------------------------------------------------------------ closure/HLL
// Levermore closure of one state (flux.py:10-61).  q = |J| / phi is kept so
// that the eigenvalues can reuse f, chi and n when their own clamps
// (phi >= EPS, f <= 1 - EPS) leave the same values.
struct Closure {
    double jn, q, chi, nx, ny, nz, iso, dir;
};

__device__ __forceinline__ double levermore_chi(double f)
{
    return div_by<3>(5.0 - 2.0 * sqrt(4.0 - 3.0 * f * f));
}

__device__ __forceinline__ Closure closure(double phi, double jx, double jy, double jz)
{
    Closure c;
    c.jn = sqrt(jx * jx + jy * jy + jz * jz);
    double f;
    if (phi > 0.0) {
        c.q = c.jn / phi;
        f = c.q;
        if (f < 0.0)
            f = 0.0;
        else if (f > 1.0)
            f = 1.0;
    } else {
        c.q = 0.0;
        f = 0.0;
    }
    c.chi = levermore_chi(f);
    if (c.jn > 0.0) {
        c.nx = jx / c.jn;
        c.ny = jy / c.jn;
        c.nz = jz / c.jn;
    } else {
        c.nx = c.ny = c.nz = 0.0;
    }
    c.iso = 0.5 * (1.0 - c.chi);
    c.dir = 0.5 * (3.0 * c.chi - 1.0);
    return c;
}

// Extreme M1 eigenvalues along an axis from (f, mu, chi) (wave_speeds.py:72-103).
__device__ __forceinline__ void eigen_core(double f, double mu, double chi, double& lmin,
                                           double& lmax)
{
    const double den = 5.0 - 3.0 * chi;
    double disc = 4.0 * f * f * mu * mu - den * (mu * mu * (3.0 * chi - 1.0) - 2.0 * (1.0 - chi));
    if (disc < 0.0)
        disc = 0.0;
    const double root = sqrt(disc);
    lmin = (2.0 * f * mu - root) / den;
    lmax = (2.0 * f * mu + root) / den;
}

// wave_speeds.py:37-103 / flux.py:63-135 for one state along an axis.
__device__ __forceinline__ void eigen(double phi, double jnormal, double jn, double& lmin,
                                      double& lmax)
{
    const double iso = 1.0 / sqrt(3.0);
    if (phi < EIG_EPS)
        phi = EIG_EPS;
    if (jn < EIG_EPS) {
        lmin = -iso;
        lmax = iso;
        return;
    }
    double f = jn / phi;
    if (f < 0.0)
        f = 0.0;
    else if (f > 1.0 - EIG_EPS)
        f = 1.0 - EIG_EPS;
    eigen_core(f, jnormal / jn, levermore_chi(f), lmin, lmax);
}

// Same eigenvalues reusing the closure of the state: with phi >= EPS and
// |J| / phi <= 1 - EPS the eigenvalue clamps are inactive and f, chi and
// mu = J_axis / |J| are the closure's values bit for bit.
template <int AXIS>
__device__ __forceinline__ void eigen_closure(double phi, const Closure& c, double jnormal,
                                              double& lmin, double& lmax)
{
    if (phi >= EIG_EPS && c.jn >= EIG_EPS && c.q <= 1.0 - EIG_EPS) {
        const double mu = AXIS == 0 ? c.nx : (AXIS == 1 ? c.ny : c.nz);
        eigen_core(c.q, mu, c.chi, lmin, lmax);
    } else {
        eigen(phi, jnormal, c.jn, lmin, lmax);
    }
}

// Physical flux along AXIS (flux.py:199-309).
template <int AXIS>
__device__ __forceinline__ void phys_flux(double phi, double jx, double jy, double jz,
                                          const Closure& c, double F[4])
{
    if (AXIS == 0) {
        F[0] = jx;
        F[1] = phi * (c.iso + c.dir * c.nx * c.nx);
        F[2] = phi * c.dir * c.nx * c.ny;
        F[3] = phi * c.dir * c.nx * c.nz;
    } else if (AXIS == 1) {
        F[0] = jy;
        F[1] = phi * c.dir * c.nx * c.ny;
        F[2] = phi * (c.iso + c.dir * c.ny * c.ny);
        F[3] = phi * c.dir * c.ny * c.nz;
    } else {
        F[0] = jz;
        F[1] = phi * c.dir * c.nx * c.nz;
        F[2] = phi * c.dir * c.ny * c.nz;
        F[3] = phi * (c.iso + c.dir * c.nz * c.nz);
    }
}

// HLL flux of one face from projected left/right states (flux.py:137-386).
template <int AXIS>
__device__ __forceinline__ void hll(const double L[4], const double R[4], int wave_mode,
                                    double F[4])
{
    const Closure cl = closure(L[0], L[1], L[2], L[3]);
    const Closure cr = closure(R[0], R[1], R[2], R[3]);
    double FL[4], FR[4];
    phys_flux<AXIS>(L[0], L[1], L[2], L[3], cl, FL);
    phys_flux<AXIS>(R[0], R[1], R[2], R[3], cr, FR);
    double sl, sr;
    if (wave_mode == 0) {
        double lminl, lmaxl, lminr, lmaxr;
        eigen_closure<AXIS>(L[0], cl, L[AXIS + 1], lminl, lmaxl);
        eigen_closure<AXIS>(R[0], cr, R[AXIS + 1], lminr, lmaxr);
        sl = lminl < lminr ? lminl : lminr;
        sr = lmaxl > lmaxr ? lmaxl : lmaxr;
        if (sl > 0.0)
            sl = 0.0;
        if (sr < 0.0)
            sr = 0.0;
    } else {
        sl = -1.0;
        sr = 1.0;
    }
    const double den = sr - sl;
    const double sp = sl * sr;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        F[c] = (sr * FL[c] - sl * FR[c] + sp * (R[c] - L[c])) / den;
}

// Face flux from the six cell states s[c][0..5] around the face (cells
// f-3 .. f+2 for the face f-1/2): WENO-Z left/right states, projection, HLL.
template <int AXIS>
__device__ __forceinline__ void face_flux(const double s[4][6], int wave_mode, double F[4])
{
    double L[4], R[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        L[c] = weno_left(s[c][0], s[c][1], s[c][2], s[c][3], s[c][4]);
        R[c] = weno_left(s[c][5], s[c][4], s[c][3], s[c][2], s[c][1]);
    }
    project(L[0], L[1], L[2], L[3]);
    project(R[0], R[1], R[2], R[3]);
    hll<AXIS>(L, R, wave_mode, F);
}

// Runtime-axis variant for the tile's upper faces (y or z per lane).
__device__ __forceinline__ void face_flux_yz(bool zaxis, const double s[4][6], int wave_mode,
                                             double F[4])
{
    if (zaxis)
        face_flux<2>(s, wave_mode, F);
    else
        face_flux<1>(s, wave_mode, F);
}

// ------------------------------------------------------------- cell loads
// Stencil input of one cell.  STAGE1 / STAGE2 load V_k and form the projected
// stage state proj(C_k V_k / S_k) (energy.py:1005-1016, 1059-1070).
template <int KIND, bool IFAC>
__device__ __forceinline__ void load_state(const StageArgs& a, const StepCoef& c, long i, long N,
                                           double u[4])
{
    u[0] = a.src[i];
    u[1] = a.src[i + N];
    u[2] = a.src[i + 2 * N];
    u[3] = a.src[i + 3 * N];
    if constexpr (KIND == STAGE1 || KIND == STAGE2) {
        const int m = a.ids[i];
        if constexpr (IFAC) {
            const double fp = KIND == STAGE1 ? c.f1p[m] : c.f2p[m];
            const double fj = KIND == STAGE1 ? c.f1j[m] : c.f2j[m];
            u[0] *= fp;
            u[1] *= fj;
            u[2] *= fj;
            u[3] *= fj;
        }
        const double S = KIND == STAGE1 ? c.S1[m] : c.S2[m];
        u[0] = u[0] / S;
        u[1] = u[1] / S;
        u[2] = u[2] / S;
        u[3] = u[3] / S;
        project(u[0], u[1], u[2], u[3]);
    }
}

// Quartic z-min ghost cell k (1, 2, 3) of one component (boundary.py:274-299).
__device__ __forceinline__ double zmin_ghost(int k, double uin, double u0, double u1, double u2,
                                             double u3)
{
    if (k == 1)
        return 5.0 * uin - (77.0 / 12.0) * u0 + (43.0 / 12.0) * u1 - (17.0 / 12.0) * u2 +
               0.25 * u3;
    if (k == 2)
        return 25.0 * uin - (505.0 / 12.0) * u0 + (335.0 / 12.0) * u1 - (145.0 / 12.0) * u2 +
               (9.0 / 4.0) * u3;
    return 75.0 * uin - (545.0 / 4.0) * u0 + (395.0 / 4.0) * u1 - (181.0 / 4.0) * u2 +
           (35.0 / 4.0) * u3;
}

// Stage update of one cell from its flux divergence rhs and its own stencil
// input w (energy.py:968-1128 stage algebra, dose quadrature, projection).
template <int KIND, bool IFAC>
__device__ __forceinline__ void stage_epilogue(const StageArgs& a, const StepCoef& sc, long i,
                                               long N, const double rhs[4], const double w[4])
{
    const int m = a.ids[i];
    const double dE = sc.dE;
    if constexpr (KIND == STAGE_RHS) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            a.out[i + c * N] = rhs[c];
    } else if constexpr (KIND == STAGE0) {
        const double S0 = sc.S0[m];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            a.out[i + c * N] = rhs[c] * dE + S0 * w[c];
    } else if constexpr (KIND == STAGE_FE) {
        // energy.py:663-722
        const double S0 = sc.S0[m], S1 = sc.S1[m];
        double u[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            u[c] = (rhs[c] * dE + S0 * w[c]) / S1;
        project(u[0], u[1], u[2], u[3]);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            a.out[i + c * N] = u[c];
        const double e = dE * w[0] * S0;
        if (a.edep)
            a.edep[i] = e;
        if (a.dose)
            a.dose[i] += e * MEV_TO_JOULE / sc.rho[m];
    } else if constexpr (KIND == STAGE1) {
        const double S0 = sc.S0[m];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double r = rhs[c];
            if constexpr (IFAC)
                r = r * (c == 0 ? sc.i1p[m] : sc.i1j[m]);
            r = r * dE;
            r = r + a.V1[i + c * N];
            r = r * 0.25;
            a.out[i + c * N] = 0.75 * (S0 * a.U[i + c * N]) + r;
        }
    } else {  // STAGE2
        const double S0 = sc.S0[m], S1 = sc.S1[m], S2 = sc.S2[m];
        double un[4], u[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            un[c] = a.U[i + c * N];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double r = rhs[c];
            if constexpr (IFAC)
                r = r * (c == 0 ? sc.i2p[m] : sc.i2j[m]);
            r = r * dE;
            r = r + a.src[i + c * N];
            r = r * (2.0 / 3.0);
            double v = (1.0 / 3.0) * (S0 * un[c]) + r;
            if constexpr (IFAC)
                v = v * (c == 0 ? sc.f1p[m] : sc.f1j[m]);
            u[c] = v / S1;
        }
        project(u[0], u[1], u[2], u[3]);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            a.U[i + c * N] = u[c];
        // Third-order dose quadrature (energy.py:1121-1128, 1617-1621).
        const double q0 = un[0] * S0;
        double u1 = a.V1[i];
        if constexpr (IFAC)
            u1 = u1 * sc.f1p[m];
        u1 = u1 / S1;
        if (u1 < 0.0)
            u1 = 0.0;
        const double q1 = u1 * S1;
        const double q2 = w[0] * S2;
        const double e = dE * ((1.0 / 6.0) * q0 + (1.0 / 6.0) * q1 + (2.0 / 3.0) * q2);
        if (a.edep)
            a.edep[i] = e;
        if (a.dose)
            a.dose[i] += e * MEV_TO_JOULE / sc.rho[m];
    }
}

// ------------------------------------------------------------ stage kernel
template <int KIND, bool IFAC>
__global__ void __launch_bounds__(NT) stage_kernel(StageArgs a)
{
    __shared__ StepCoef sc;
    __shared__ double P[4][PY][PZ];
    __shared__ double Fy[4][TY + 1][TZ];
    __shared__ double Fz[4][TY][TZ + 1];

    const int tid = threadIdx.x;
    {
        const double* src = reinterpret_cast<const double*>(a.coef);
        double* dst = reinterpret_cast<double*>(&sc);
        for (int k = tid; k < int(sizeof(StepCoef) / sizeof(double)); k += NT)
            dst[k] = src[k];
    }
    __syncthreads();

    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long N = long(nx) * ny * nz;
    const int tz = tid % TZ, ty = tid / TZ;
    const int z0 = blockIdx.x * TZ, y0 = blockIdx.y * TY, x0 = blockIdx.z * TX;
    const int z = z0 + tz, y = y0 + ty;
    const int zc = min(z, nz - 1), yc = min(y, ny - 1);
    const bool valid = z < nz && y < ny;
    const int x1 = min(x0 + TX, nx);
    const int wave = a.wave_mode;

    auto cidx = [&](int xx, int yy, int zz) -> long { return (long(xx) * ny + yy) * nz + zz; };
    auto clampx = [&](int xx) { return xx < 0 ? 0 : (xx >= nx ? nx - 1 : xx); };

#if M1_EXACT
    const double dx = a.g.dx, dy = a.g.dy, dz = a.g.dz;
#else
    const double idx = 1.0 / a.g.dx, idy = 1.0 / a.g.dy, idz = 1.0 / a.g.dz;
#endif

    // Halo cell of this thread in the padded plane (hy, hz), if any.
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

    // x-window: W[k] holds plane (x - 2 + k) of this thread's column.
    double W[6][4];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        load_state<KIND, IFAC>(a, sc, cidx(clampx(x0 - 3 + k), yc, zc), N, W[k]);
    double Flo[4];
    {
        double s[4][6];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                s[c][k] = W[k][c];
        face_flux<0>(s, wave, Flo);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            W[k][c] = W[k + 1][c];

    for (int x = x0; x < x1; ++x) {
        load_state<KIND, IFAC>(a, sc, cidx(clampx(x + 3), yc, zc), N, W[5]);

        // Plane x into LDS: own cell and one halo cell per thread.
#pragma unroll
        for (int c = 0; c < 4; ++c)
            P[c][ty + 3][tz + 3] = W[2][c];
        if (hy >= 0) {
            double u[4];
            if (hgz_raw < 0) {
                // z-min inflow ghost from the boundary state and cells 0..3.
                double q[4][4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    load_state<KIND, IFAC>(a, sc, cidx(x, hgy, k), N, q[k]);
                const long b = long(x) * ny + hgy;
                const long m2 = long(nx) * ny;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    u[c] = zmin_ghost(-hgz_raw, a.bnd[c * m2 + b], q[0][c], q[1][c], q[2][c],
                                      q[3][c]);
            } else {
                load_state<KIND, IFAC>(a, sc, cidx(x, hgy, hgz), N, u);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
                P[c][hy][hz] = u[c];
        }

        // x face ahead of the march (planes x-2 .. x+3).
        double Fhi[4];
        {
            double s[4][6];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    s[c][k] = W[k][c];
            face_flux<0>(s, wave, Fhi);
        }
        __syncthreads();

        // Lower z and y faces of this cell.
        {
            double s[4][6], F[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    s[c][k] = P[c][ty + 3][tz + k];
            face_flux<2>(s, wave, F);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Fz[c][ty][tz] = F[c];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    s[c][k] = P[c][ty + k][tz + 3];
            face_flux<1>(s, wave, F);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Fy[c][ty][tz] = F[c];
        }
        // Upper faces of the tile: z face of each row, y face of each column.
        if (tid < TY + TZ) {
            const bool zaxis = tid < TY;
            const int row = zaxis ? tid : 0;
            const int col = zaxis ? 0 : tid - TY;
            double s[4][6], F[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    s[c][k] = zaxis ? P[c][row + 3][TZ + k] : P[c][TY + k][col + 3];
            face_flux_yz(zaxis, s, wave, F);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (zaxis)
                    Fz[c][row][TZ] = F[c];
                else
                    Fy[c][TY][col] = F[c];
            }
        }
        __syncthreads();

        if (valid) {
            const long i = cidx(x, y, z);
            double rhs[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#if M1_EXACT
                const double ax = (Fhi[c] - Flo[c]) / dx;
                const double ay = (Fy[c][ty + 1][tz] - Fy[c][ty][tz]) / dy;
                const double az = (Fz[c][ty][tz + 1] - Fz[c][ty][tz]) / dz;
#else
                const double ax = (Fhi[c] - Flo[c]) * idx;
                const double ay = (Fy[c][ty + 1][tz] - Fy[c][ty][tz]) * idy;
                const double az = (Fz[c][ty][tz + 1] - Fz[c][ty][tz]) * idz;
#endif
                rhs[c] = -(ax + ay + az);
            }
            stage_epilogue<KIND, IFAC>(a, sc, i, N, rhs, W[2]);
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
}

// ------------------------------------------------------ split stage kernels
// The same stage as stage_kernel in three passes with one thread per cell or
// per face and no tile state held across a march: (1) stages 1 and 2 form
// their stencil input proj(C_k V_k / S_k) once per cell into `win`; (2) one
// kernel per axis computes every face flux of that axis from the six cells
// around the face (ghosts as in stage_kernel: zero gradient, quartic z-min
// inflow) into fx / fy / fz; (3) the update kernel takes the divergence and
// runs the stage epilogue.  Each pass needs a fraction of stage_kernel's
// registers, so the double-precision WENO-Z / closure / HLL arithmetic runs
// at several waves per SIMD instead of one; the flux arrays cost HBM traffic
// that this VALU-bound arithmetic hides.  The arithmetic of every face and
// cell is stage_kernel's, so both forms give the same bits.
constexpr int SPLIT_THREADS = 256;
#ifndef M1_F64_SPLIT_WAVES
#define M1_F64_SPLIT_WAVES 2
#endif

template <int KIND, bool IFAC>
__global__ void __launch_bounds__(SPLIT_THREADS) input_kernel(StageArgs a)
{
    const long N = long(a.g.nx) * a.g.ny * a.g.nz;
    const long i = long(blockIdx.x) * SPLIT_THREADS + threadIdx.x;
    if (i >= N)
        return;
    double u[4];
    load_state<KIND, IFAC>(a, *a.coef, i, N, u);
#pragma unroll
    for (int c = 0; c < 4; ++c)
        a.win[i + c * N] = u[c];
}

// Faces of AXIS: face f (0..n) lies between cells f-1 and f of that axis.
template <int AXIS>
__global__ void __launch_bounds__(SPLIT_THREADS, M1_F64_SPLIT_WAVES)
face_kernel(StageArgs a, const double* __restrict__ in, double* __restrict__ flux)
{
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long N = long(nx) * ny * nz;
    const int fny = ny + (AXIS == 1), fnz = nz + (AXIS == 2);
    const long nf = long(nx + (AXIS == 0)) * fny * fnz;
    const long t = long(blockIdx.x) * SPLIT_THREADS + threadIdx.x;
    if (t >= nf)
        return;
    const int z = int(t % fnz);
    const long r = t / fnz;
    const int y = int(r % fny);
    const int x = int(r / fny);
    auto cidx = [&](int xx, int yy, int zz) -> long { return (long(xx) * ny + yy) * nz + zz; };
    double s[4][6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if constexpr (AXIS == 0) {
            const int xx = min(max(x - 3 + k, 0), nx - 1);
            const long i = cidx(xx, y, z);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                s[c][k] = in[i + c * N];
        } else if constexpr (AXIS == 1) {
            const int yy = min(max(y - 3 + k, 0), ny - 1);
            const long i = cidx(x, yy, z);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                s[c][k] = in[i + c * N];
        } else {
            const int zz = z - 3 + k;
            if (zz < 0) {
                // z-min inflow ghost from the boundary state and cells 0..3.
                const long b = long(x) * ny + y;
                const long m2 = long(nx) * ny;
                const long i0 = cidx(x, y, 0);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    s[c][k] = zmin_ghost(-zz, a.bnd[c * m2 + b], in[i0 + c * N],
                                         in[i0 + 1 + c * N], in[i0 + 2 + c * N],
                                         in[i0 + 3 + c * N]);
            } else {
                const long i = cidx(x, y, min(zz, nz - 1));
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    s[c][k] = in[i + c * N];
            }
        }
    }
    double F[4];
    face_flux<AXIS>(s, a.wave_mode, F);
#pragma unroll
    for (int c = 0; c < 4; ++c)
        flux[t + c * nf] = F[c];
}

template <int KIND, bool IFAC>
__global__ void __launch_bounds__(SPLIT_THREADS) update_kernel(StageArgs a,
                                                               const double* __restrict__ in)
{
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long N = long(nx) * ny * nz;
    const long i = long(blockIdx.x) * SPLIT_THREADS + threadIdx.x;
    if (i >= N)
        return;
    const int z = int(i % nz);
    const int y = int((i / nz) % ny);
    const int x = int(i / (long(ny) * nz));
    const long nfx = long(nx + 1) * ny * nz, nfy = long(nx) * (ny + 1) * nz,
               nfz = long(nx) * ny * (nz + 1);
    const long ix = (long(x) * ny + y) * nz + z;            // face x (lower) of the cell
    const long iy = (long(x) * (ny + 1) + y) * nz + z;      // face y (lower)
    const long iz = (long(x) * ny + y) * (nz + 1) + z;      // face z (lower)
    const long sx = long(ny) * nz;
#if !M1_EXACT
    const double idx = 1.0 / a.g.dx, idy = 1.0 / a.g.dy, idz = 1.0 / a.g.dz;
#endif
    double rhs[4], w[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double fxl = a.fx[ix + c * nfx], fxh = a.fx[ix + sx + c * nfx];
        const double fyl = a.fy[iy + c * nfy], fyh = a.fy[iy + nz + c * nfy];
        const double fzl = a.fz[iz + c * nfz], fzh = a.fz[iz + 1 + c * nfz];
#if M1_EXACT
        const double ax = (fxh - fxl) / a.g.dx;
        const double ay = (fyh - fyl) / a.g.dy;
        const double az = (fzh - fzl) / a.g.dz;
#else
        const double ax = (fxh - fxl) * idx;
        const double ay = (fyh - fyl) * idy;
        const double az = (fzh - fzl) * idz;
#endif
        rhs[c] = -(ax + ay + az);
        w[c] = in[i + c * N];
    }
    stage_epilogue<KIND, IFAC>(a, *a.coef, i, N, rhs, w);
}

template <int KIND, bool IFAC>
hipError_t launch_split(const StageArgs& a, hipStream_t s)
{
    const long N = long(a.g.nx) * a.g.ny * a.g.nz;
    const double* in = a.src;
    if constexpr (KIND == STAGE1 || KIND == STAGE2) {
        hipLaunchKernelGGL((input_kernel<KIND, IFAC>), dim3(unsigned((N + 255) / 256)),
                           dim3(SPLIT_THREADS), 0, s, a);
        in = a.win;
    }
    const long nfx = long(a.g.nx + 1) * a.g.ny * a.g.nz;
    const long nfy = long(a.g.nx) * (a.g.ny + 1) * a.g.nz;
    const long nfz = long(a.g.nx) * a.g.ny * (a.g.nz + 1);
    hipLaunchKernelGGL(face_kernel<0>, dim3(unsigned((nfx + 255) / 256)), dim3(SPLIT_THREADS), 0,
                       s, a, in, a.fx);
    hipLaunchKernelGGL(face_kernel<1>, dim3(unsigned((nfy + 255) / 256)), dim3(SPLIT_THREADS), 0,
                       s, a, in, a.fy);
    hipLaunchKernelGGL(face_kernel<2>, dim3(unsigned((nfz + 255) / 256)), dim3(SPLIT_THREADS), 0,
                       s, a, in, a.fz);
    hipLaunchKernelGGL((update_kernel<KIND, IFAC>), dim3(unsigned((N + 255) / 256)),
                       dim3(SPLIT_THREADS), 0, s, a, in);
    return hipGetLastError();
}

// -------------------------------------------------------------- CFL kernel
// compute_energy_step (energy.py:539-660): max over cells of
// sum_axis a_axis / (S dx_axis), with the inflow speeds folded into z = 0.
constexpr int CFL_THREADS = 256;

__global__ void __launch_bounds__(CFL_THREADS) cfl_kernel(CflArgs a)
{
    __shared__ double S0[MAXMAT];
    __shared__ double red[CFL_THREADS / 64];
    if (threadIdx.x < MAXMAT)
        S0[threadIdx.x] = a.coef->S0[threadIdx.x];
    __syncthreads();
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long N = long(nx) * ny * nz;
    const double ed = a.coef->ed[0], rf = a.coef->rf;
    double mx = 0.0;
    bool bad = false;
    for (long i = long(blockIdx.x) * CFL_THREADS + threadIdx.x; i < N;
         i += long(gridDim.x) * CFL_THREADS) {
        const double s = S0[a.ids[i]];
        double v;
        if (a.wave_mode == 1) {
            v = 1.0 / (s * a.g.dx) + 1.0 / (s * a.g.dy) + 1.0 / (s * a.g.dz);
        } else {
            const double phi = a.U[i], jx = a.U[i + N], jy = a.U[i + 2 * N], jz = a.U[i + 3 * N];
            const double jn = sqrt(jx * jx + jy * jy + jz * jz);
            double ax, ay, az, lo, hi;
            if (jn < EIG_EPS) {
                ax = ay = az = 1.0 / sqrt(3.0);
            } else {
                // f and chi are the same for the three axes.
                double f = jn / (phi < EIG_EPS ? EIG_EPS : phi);
                if (f < 0.0)
                    f = 0.0;
                else if (f > 1.0 - EIG_EPS)
                    f = 1.0 - EIG_EPS;
                const double chi = levermore_chi(f);
                eigen_core(f, jx / jn, chi, lo, hi);
                ax = fmax(fabs(lo), fabs(hi));
                eigen_core(f, jy / jn, chi, lo, hi);
                ay = fmax(fabs(lo), fabs(hi));
                eigen_core(f, jz / jn, chi, lo, hi);
                az = fmax(fabs(lo), fabs(hi));
            }
            if (i % nz == 0) {
                const double pin = a.spatial[i / nz] * ed;
                const double jin = rf * pin;
                const double jnin = sqrt(0.0 * 0.0 + 0.0 * 0.0 + jin * jin);
                eigen(pin, jin, jnin, lo, hi);
                az = fmax(az, fmax(fabs(lo), fabs(hi)));
            }
            if (a.act) {
                // energy.py:624-630: an axis's speed counts only within
                // act_radius cells (along that axis) of an active cell.
                const long sx = long(ny) * nz;
                const long ix = i / sx, iy = (i / nz) % ny, iz = i % nz;
                bool mx_on = false, my_on = false, mz_on = false;
                for (int d = -a.act_radius; d <= a.act_radius; ++d) {
                    if (ix + d >= 0 && ix + d < nx)
                        mx_on |= a.act[i + d * sx] != 0;
                    if (iy + d >= 0 && iy + d < ny)
                        my_on |= a.act[i + d * nz] != 0;
                    if (iz + d >= 0 && iz + d < nz)
                        mz_on |= a.act[i + d] != 0;
                }
                ax = mx_on ? ax : 0.0;
                ay = my_on ? ay : 0.0;
                az = mz_on ? az : 0.0;
            }
            v = ax / (s * a.g.dx) + ay / (s * a.g.dy) + az / (s * a.g.dz);
        }
        if (!(v == v))
            bad = true;
        else
            mx = fmax(mx, v);
    }
    if (bad)
        mx = __longlong_as_double(0x7ff8000000000000LL);
    // wavefront then workgroup max; NaN propagates as the largest bit pattern.
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
        for (int w = 0; w < CFL_THREADS / 64; ++w) {
            const unsigned long long o = __double_as_longlong(red[w]);
            b = o > b ? o : b;
        }
        atomicMax(a.out, b);
    }
}

// --------------------------------------------------------- boundary kernel
// Stage boundary traces at z-min: U_in(E) for stage 0, and the RK-consistent
// stage-1/2 states of _ssprk3_stage_inflow_states (energy.py:107-230) or
// _if_ssprk3_stage_inflow_states (energy.py:348-455).
__global__ void boundary_kernel(BoundaryArgs a)
{
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long m2 = long(nx) * ny;
    const long b = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= m2)
        return;
    const StepCoef& c = *a.coef;
    const int m = a.ids[b * nz];
    const double sp = a.spatial[b];
    double* o0 = a.bnd;
    double* o1 = a.bnd + 4 * m2;
    double* o2 = a.bnd + 8 * m2;
    // stage 0: inflow.state(E) (boundary.py:101-112)
    const double p0 = sp * c.ed[0];
    o0[b] = p0;
    o0[m2 + b] = 0.0;
    o0[2 * m2 + b] = 0.0;
    o0[3 * m2 + b] = c.rf * p0;
    // conserved traces B(E_k) = S(E_k) U_in(E_k), k = 0, 1/3, 2/3, 1
    const double Sk[4] = {c.S0[m], c.S13[m], c.S23[m], c.S1[m]};
    double B[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double p = sp * c.ed[k];
        const double s = Sk[k];
        B[k][0] = s * p;
        B[k][1] = s * 0.0;
        B[k][2] = s * 0.0;
        B[k][3] = s * (c.rf * p);
    }
    if (a.integrating_factor) {
        const double ip[3] = {c.i13p[m], c.i23p[m], c.i1p[m]};
        const double ij[3] = {c.i13j[m], c.i23j[m], c.i1j[m]};
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            B[k][0] *= ip[k - 1];
            B[k][1] *= ij[k - 1];
            B[k][2] *= ij[k - 1];
            B[k][3] *= ij[k - 1];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v1 = 0.5 * (-9.0 * B[0][q] + 18.0 * B[1][q] - 9.0 * B[2][q] + 2.0 * B[3][q]);
        double v2 = 0.25 * (11.0 * B[0][q] - 27.0 * B[1][q] + 27.0 * B[2][q] - 7.0 * B[3][q]);
        if (a.integrating_factor) {
            v1 = v1 * (q == 0 ? c.f1p[m] : c.f1j[m]);
            v2 = v2 * (q == 0 ? c.f2p[m] : c.f2j[m]);
        }
        o1[q * m2 + b] = v1 / c.S1[m];
        o2[q * m2 + b] = v2 / c.S2[m];
    }
}

// Strang half-step source flows (energy.py:1168-1191): nuclear damping of all
// moments, then scattering damping of J.
__global__ void scale_kernel(ScaleArgs a)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    const int m = a.ids[i];
    const StepCoef& c = *a.coef;
    if (a.nuclear) {
        const double f = c.hn[m];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            a.U[i + q * a.n] *= f;
    }
    if (a.scattering) {
        const double f = c.hs[m];
#pragma unroll
        for (int q = 1; q < 4; ++q)
            a.U[i + q * a.n] *= f;
    }
}

}  // namespace

hipError_t launch_stage(int kind, bool ifac, const StageArgs& a, hipStream_t s)
{
    if (a.fx && a.fy && a.fz && a.win) {
        switch (kind) {
        case STAGE0:
            return launch_split<STAGE0, false>(a, s);
        case STAGE1:
            return ifac ? launch_split<STAGE1, true>(a, s) : launch_split<STAGE1, false>(a, s);
        case STAGE2:
            return ifac ? launch_split<STAGE2, true>(a, s) : launch_split<STAGE2, false>(a, s);
        case STAGE_RHS:
            return launch_split<STAGE_RHS, false>(a, s);
        case STAGE_FE:
            return launch_split<STAGE_FE, false>(a, s);
        default:
            return hipErrorInvalidValue;
        }
    }
    const dim3 grid((a.g.nz + TZ - 1) / TZ, (a.g.ny + TY - 1) / TY, (a.g.nx + TX - 1) / TX);
    const dim3 block(NT);
#define M1_LAUNCH(K, I) hipLaunchKernelGGL((stage_kernel<K, I>), grid, block, 0, s, a)
    switch (kind) {
    case STAGE0:
        M1_LAUNCH(STAGE0, false);
        break;
    case STAGE1:
        if (ifac)
            M1_LAUNCH(STAGE1, true);
        else
            M1_LAUNCH(STAGE1, false);
        break;
    case STAGE2:
        if (ifac)
            M1_LAUNCH(STAGE2, true);
        else
            M1_LAUNCH(STAGE2, false);
        break;
    case STAGE_RHS:
        M1_LAUNCH(STAGE_RHS, false);
        break;
    case STAGE_FE:
        M1_LAUNCH(STAGE_FE, false);
        break;
    default:
        return hipErrorInvalidValue;
    }
#undef M1_LAUNCH
    return hipGetLastError();
}

hipError_t launch_cfl(const CflArgs& a, hipStream_t s)
{
    const long N = long(a.g.nx) * a.g.ny * a.g.nz;
    long blocks = (N + CFL_THREADS - 1) / CFL_THREADS;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(cfl_kernel, dim3(unsigned(blocks)), dim3(CFL_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_boundary(const BoundaryArgs& a, hipStream_t s)
{
    const long m2 = long(a.g.nx) * a.g.ny;
    hipLaunchKernelGGL(boundary_kernel, dim3(unsigned((m2 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_scale(const ScaleArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(scale_kernel, dim3(unsigned((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace M1_NS
}  // namespace m1
