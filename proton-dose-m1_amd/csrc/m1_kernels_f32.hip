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

template <int KIND, bool IFAC, bool D>
__global__ void __launch_bounds__(NT, M1_STAGE_WAVES) stage_kernel(StageArgs32 a)
{
    const long tile = (long(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    if (a.compute && !a.compute[tile])
        return;  // whole tile and its 9-cell neighbourhood are vacuum
    TileView v{a.g,
               a.src,
               a.U,
               a.V1,
               a.out,
               a.dose_hi,
               a.dose_lo,
               a.edep,
               a.ids,
               a.bnd,
               a.coef,
               a.tile_max ? a.tile_max + tile : nullptr,
               nullptr,
               a.tile_face ? a.tile_face + FACE_SLOTS * tile : nullptr,
               a.wave_mode,
               int(blockIdx.z),
               int(blockIdx.y),
               int(blockIdx.x)};
    stage_tile<KIND, IFAC, D>(v);
}

// ---------------------------------------------------------------- CFL
constexpr int CFL_THREADS = 256;

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
    double mx = 0.0;
    bool bad = false;
    for (long i = long(blockIdx.x) * CFL_THREADS + threadIdx.x; i < N;
         i += long(gridDim.x) * CFL_THREADS) {
        const bool zmin = i % nz == 0;
        const double pin = zmin ? a.spatial[i / nz] * pin_scale : 0.0;
        const double v = cfl_cell(a.U, i, N, S0[a.ids[i]], a.g, a.wave_mode, eps, zmin, pin, rf);
        if (!(v == v))
            bad = true;
        else
            mx = fmax(mx, v);
    }
    block_max_to(mx, bad, red, a.out);
}

// Stage boundary traces (energy.py:107-230, 348-455) in double, stored
// scaled by 1/alpha in float.
__global__ void boundary_kernel(BoundaryArgs32 a)
{
    const long m2 = long(a.g.nx) * a.g.ny;
    const long b = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= m2)
        return;
    boundary_column(*a.coef, a.ids[b * a.g.nz], a.spatial[b], b, m2, a.integrating_factor != 0,
                    a.bnd);
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

// Next-step tile set: a tile is evaluated when it holds phi above the
// threshold after this step, when a cell within TILE_REACH cells of it in a
// neighbouring tile does (bounded by that tile's face slabs), or when it is a
// z = 0 tile while the inflow spectrum is on.
__global__ void tile_mask_kernel(TileMaskArgs a)
{
    const int n = a.ntx * a.nty * a.ntz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int tz = t % a.ntz, ty = (t / a.ntz) % a.nty, tx = t / (a.ntz * a.nty);
    bool on = (a.inflow_on && tz == 0) || (a.computed[t] && a.tile_max[t] > a.threshold);
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
                if (zz < 0 || zz >= a.ntz || (dx == 0 && dy == 0 && dz == 0))
                    continue;
                const int u = (xx * a.nty + yy) * a.ntz + zz;
                on = a.computed[u] &&
                     facing_bound(a.tile_face + FACE_SLOTS * long(u), dx, dy, dz, 0) > a.threshold;
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
#define M1_LAUNCH(K, I)                                                                   \
    do {                                                                                  \
        if (a.f64_closure)                                                                \
            hipLaunchKernelGGL((stage_kernel<K, I, true>), grid, block, 0, s, a);         \
        else                                                                              \
            hipLaunchKernelGGL((stage_kernel<K, I, false>), grid, block, 0, s, a);        \
    } while (0)
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
