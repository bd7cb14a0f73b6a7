// FP32-bulk march kernels for gfx950.
//
// Same tiling and stage structure as m1_kernels.hip (one fused kernel per
// SSPRK3 stage: WENO-Z in three axes, HLL, divergence, stage algebra), with
// the arithmetic split by what each part needs:
//   * state, stage variables, WENO-Z reconstruction, divergence and stage
//     combinations in float, on the state normalised by alpha (the peak
//     inflow fluence of the march), so that phi is O(1) and the smoothness
//     indicators (O(phi^2)) stay far from float overflow.  The M1 system is
//     positively homogeneous in U, so the march is unchanged when the
//     absolute thresholds are rescaled with it: the WENO epsilon by
//     1/alpha^2 and the eigenvalue floors by 1/alpha;
//   * realizability projection, closure, eigenvalues and the HLL combination
//     in double, because the eigenvalue discriminant cancels and 1 - chi loses
//     its digits as |J|/phi -> 1 (flux.py:108-117), and the projection
//     margin 1e-10 is below float resolution;
//   * inflow ghost extrapolation (coefficients up to 136) in double;
//   * dose accumulated in a compensated float pair (TwoSum), read out as
//     double(hi) + double(lo).
#include "m1_f32_common.h"

namespace m1 {
namespace f32 {
namespace {

using namespace detail;

template <int KIND, bool IFAC>
__global__ void __launch_bounds__(NT) stage_kernel(StageArgs32 a)
{
    const long tile = (long(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    if (a.compute && !a.compute[tile])
        return;  // whole tile and its 9-cell neighbourhood are vacuum
    TileView v{a.g,      a.src,       a.U,   a.V1,  a.out,  a.dose_hi,
               a.dose_lo, a.edep,     a.ids, a.bnd, a.coef, a.tile_max ? a.tile_max + tile : nullptr,
               a.wave_mode, int(blockIdx.z), int(blockIdx.y), int(blockIdx.x)};
    stage_tile<KIND, IFAC>(v);
}

// ---------------------------------------------------------------- CFL
constexpr int CFL_THREADS = 256;

__device__ __forceinline__ double amax(double lo, double hi) { return fmax(fabs(lo), fabs(hi)); }

__global__ void __launch_bounds__(CFL_THREADS) cfl_kernel(CflArgs32 a)
{
    __shared__ double S0[MAXMAT];
    __shared__ double red[CFL_THREADS / 64];
    if (threadIdx.x < MAXMAT)
        S0[threadIdx.x] = a.coef->S0[threadIdx.x];
    __syncthreads();
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long N = long(nx) * ny * nz;
    const double eps = a.coef->eig_eps;
    const double pin_scale = a.coef->ed[0] * a.coef->inv_alpha, rf = a.coef->rf;
    const double iso = 0.57735026918962576;
    double mx = 0.0;
    bool bad = false;
    for (long i = long(blockIdx.x) * CFL_THREADS + threadIdx.x; i < N;
         i += long(gridDim.x) * CFL_THREADS) {
        const double s = S0[a.ids[i]];
        double v;
        if (a.wave_mode == 1) {
            v = 1.0 / (s * a.g.dx) + 1.0 / (s * a.g.dy) + 1.0 / (s * a.g.dz);
        } else {
            const double phi = a.U[i];
            const double J[3] = {a.U[i + N], a.U[i + 2 * N], a.U[i + 3 * N]};
            const double jn = sqrt(J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
            Closure c = closure(phi, J, jn);
            double lo, hi, ax, ay, az;
            eigen<0>(phi, c, eps, lo, hi);
            ax = amax(lo, hi);
            eigen<1>(phi, c, eps, lo, hi);
            ay = amax(lo, hi);
            eigen<2>(phi, c, eps, lo, hi);
            az = amax(lo, hi);
            if (i % nz == 0) {
                const double p = a.spatial[i / nz] * pin_scale;
                const double Jin[3] = {0.0, 0.0, rf * p};
                const Closure ci = closure(p, Jin, fabs(Jin[2]));
                eigen<2>(p, ci, eps, lo, hi);
                az = fmax(az, amax(lo, hi));
            }
            (void)iso;
            v = ax / (s * a.g.dx) + ay / (s * a.g.dy) + az / (s * a.g.dz);
        }
        if (!(v == v))
            bad = true;
        else
            mx = fmax(mx, v);
    }
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
        for (int w = 0; w < CFL_THREADS / 64; ++w) {
            const unsigned long long o = __double_as_longlong(red[w]);
            b = o > b ? o : b;
        }
        atomicMax(a.out, b);
    }
}

// Stage boundary traces (energy.py:107-230, 348-455) in double, stored
// scaled by 1/alpha in float.
__global__ void boundary_kernel(BoundaryArgs32 a)
{
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long m2 = long(nx) * ny;
    const long b = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= m2)
        return;
    const StepCoef& c = *a.coef;
    const int m = a.ids[b * nz];
    const double sp = a.spatial[b];
    const double ia = c.inv_alpha;
    float* o0 = a.bnd;
    float* o1 = a.bnd + 4 * m2;
    float* o2 = a.bnd + 8 * m2;
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
    if (a.integrating_factor) {
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
        if (a.integrating_factor) {
            v1 *= (q == 0 ? c.f1p[m] : c.f1j[m]);
            v2 *= (q == 0 ? c.f2p[m] : c.f2j[m]);
        }
        o1[q * m2 + b] = float(v1 / c.S1[m] * ia);
        o2[q * m2 + b] = float(v2 / c.S2[m] * ia);
    }
}

__global__ void finalize_dose_kernel(const float* hi, const float* lo, double* dose, long n)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        dose[i] = double(hi[i]) + double(lo[i]);
}

__global__ void widen_kernel(const float* U, double alpha, double* out, long n)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = alpha * double(U[i]);
}

__global__ void narrow_kernel(const double* U, double inv_alpha, float* out, long n)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = float(U[i] * inv_alpha);
}

__global__ void scale_kernel(long n, float* U, const uint8_t* ids, const StepCoef* coef,
                             int nuclear, int scattering)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int m = ids[i];
    if (nuclear) {
        const float f = float(coef->hn[m]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            U[i + q * n] *= f;
    }
    if (scattering) {
        const float f = float(coef->hs[m]);
#pragma unroll
        for (int q = 1; q < 4; ++q)
            U[i + q * n] *= f;
    }
}

// Next-step tile set: a tile is evaluated when any tile within the dilation
// radius holds phi above the threshold after this step, or when it is a
// z = 0 tile while the inflow spectrum is on.
__global__ void tile_mask_kernel(TileMaskArgs a)
{
    const int n = a.ntx * a.nty * a.ntz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int tz = t % a.ntz, ty = (t / a.ntz) % a.nty, tx = t / (a.ntz * a.nty);
    bool on = a.inflow_on && tz == 0;
    for (int dx = -a.rx; dx <= a.rx && !on; ++dx) {
        const int xx = tx + dx;
        if (xx < 0 || xx >= a.ntx)
            continue;
        for (int dy = -a.ry; dy <= a.ry && !on; ++dy) {
            const int yy = ty + dy;
            if (yy < 0 || yy >= a.nty)
                continue;
            for (int dz = -a.rz; dz <= a.rz && !on; ++dz) {
                const int zz = tz + dz;
                if (zz < 0 || zz >= a.ntz)
                    continue;
                const int u = (xx * a.nty + yy) * a.ntz + zz;
                on = a.computed[u] && a.tile_max[u] > a.threshold;
            }
        }
    }
    a.next[t] = on ? 1 : 0;
    if (on)
        atomicAdd(a.n_active, 1);
}

// Tiles that leave the evaluated set become exact vacuum (U = V1 = V2 = 0).
__global__ void tile_zero_kernel(TileZeroArgs a)
{
    const long t = (long(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    if (!(a.computed[t] && !a.next[t]))
        return;
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const long N = long(nx) * ny * nz;
    const int z = blockIdx.x * TZ + threadIdx.x % TZ;
    const int y = blockIdx.y * TY + threadIdx.x / TZ;
    if (z >= nz || y >= ny)
        return;
    for (int x = blockIdx.z * TX; x < min(int(blockIdx.z) * TX + TX, nx); ++x) {
        const long i = (long(x) * ny + y) * nz + z;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            a.U[i + c * N] = 0.f;
            a.V1[i + c * N] = 0.f;
            a.V2[i + c * N] = 0.f;
        }
    }
}

}  // namespace

hipError_t launch_tile_mask(const TileMaskArgs& a, hipStream_t s)
{
    const int n = a.ntx * a.nty * a.ntz;
    hipLaunchKernelGGL(tile_mask_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tile_zero(const TileZeroArgs& a, hipStream_t s)
{
    const dim3 grid((a.g.nz + TZ - 1) / TZ, (a.g.ny + TY - 1) / TY, (a.g.nx + TX - 1) / TX);
    hipLaunchKernelGGL(tile_zero_kernel, grid, dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_stage(int kind, bool ifac, const StageArgs32& a, hipStream_t s)
{
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

hipError_t launch_cfl(const CflArgs32& a, hipStream_t s)
{
    const long N = long(a.g.nx) * a.g.ny * a.g.nz;
    long blocks = (N + CFL_THREADS - 1) / CFL_THREADS;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(cfl_kernel, dim3(unsigned(blocks)), dim3(CFL_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_boundary(const BoundaryArgs32& a, hipStream_t s)
{
    const long m2 = long(a.g.nx) * a.g.ny;
    hipLaunchKernelGGL(boundary_kernel, dim3(unsigned((m2 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_finalize_dose(const float* hi, const float* lo, double* dose, long n,
                                hipStream_t s)
{
    hipLaunchKernelGGL(finalize_dose_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, hi,
                       lo, dose, n);
    return hipGetLastError();
}

hipError_t launch_widen_state(const float* U, double alpha, double* out, long n, hipStream_t s)
{
    hipLaunchKernelGGL(widen_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, U, alpha,
                       out, n);
    return hipGetLastError();
}

hipError_t launch_narrow_state(const double* U, double inv_alpha, float* out, long n,
                               hipStream_t s)
{
    hipLaunchKernelGGL(narrow_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, U,
                       inv_alpha, out, n);
    return hipGetLastError();
}

hipError_t launch_scale(long n, float* U, const uint8_t* ids, const StepCoef* coef, int nuclear,
                        int scattering, hipStream_t s)
{
    hipLaunchKernelGGL(scale_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, U, ids,
                       coef, nuclear, scattering);
    return hipGetLastError();
}

}  // namespace f32
}  // namespace m1
