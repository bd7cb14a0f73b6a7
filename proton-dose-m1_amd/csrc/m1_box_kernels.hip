// Kernels of the moving active domain (numerics/moving_active_domain.py:35-253,
// numerics/energy.py:1262-1494): copies of the state between boxes of the
// grid, the box's material ids and inflow profile, the scatter of a box's
// dose into the full-grid dose, and the reductions that find the significant
// cells of the box state (and the activity of the vacuum-aware CFL).
//
// All of it is data movement and comparison: the only arithmetic (the
// threshold tol * scale, the inflow phi spatial * ed and the dose sum) is one
// separately rounded operation each, as in the reference.
#include <hip/hip_runtime.h>

#include <climits>

#include "m1_device.h"

namespace m1 {
namespace box {
namespace {

constexpr int THREADS = 256;

unsigned grid_for(long n)
{
    long b = (n + THREADS - 1) / THREADS;
    if (b > 16384)
        b = 16384;
    return unsigned(b < 1 ? 1 : b);
}

__device__ inline unsigned long long wave_max(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ inline int wave_min_i(int v)
{
    for (int off = 32; off > 0; off >>= 1)
        v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ inline int wave_max_i(int v)
{
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ void __launch_bounds__(THREADS)
copy_kernel(const double* __restrict__ src, CellBox sb, double* __restrict__ dst, CellBox db)
{
    const int nx = db.i1 - db.i0, ny = db.j1 - db.j0, nz = db.k1 - db.k0;
    const int sny = sb.j1 - sb.j0, snz = sb.k1 - sb.k0;
    const long n = long(nx) * ny * nz;
    const long sn = long(sb.i1 - sb.i0) * sny * snz;
    for (long t = long(blockIdx.x) * THREADS + threadIdx.x; t < 4 * n;
         t += long(gridDim.x) * THREADS) {
        const long c = t / n, l = t - c * n;
        const int k = int(l % nz), j = int((l / nz) % ny), i = int(l / (long(ny) * nz));
        const int gi = db.i0 + i, gj = db.j0 + j, gk = db.k0 + k;
        double v = 0.0;
        if (gi >= sb.i0 && gi < sb.i1 && gj >= sb.j0 && gj < sb.j1 && gk >= sb.k0 && gk < sb.k1)
            v = src[c * sn + (long(gi - sb.i0) * sny + (gj - sb.j0)) * snz + (gk - sb.k0)];
        dst[t] = v;
    }
}

__global__ void __launch_bounds__(THREADS)
gather_kernel(const uint8_t* __restrict__ ids, const double* __restrict__ spatial, int fny,
              int fnz, CellBox b, uint8_t* __restrict__ box_ids, double* __restrict__ box_spatial)
{
    const int nx = b.i1 - b.i0, ny = b.j1 - b.j0, nz = b.k1 - b.k0;
    const long n = long(nx) * ny * nz;
    for (long l = long(blockIdx.x) * THREADS + threadIdx.x; l < n;
         l += long(gridDim.x) * THREADS) {
        const int k = int(l % nz), j = int((l / nz) % ny), i = int(l / (long(ny) * nz));
        const long g = (long(b.i0 + i) * fny + (b.j0 + j)) * fnz + (b.k0 + k);
        box_ids[l] = ids[g];
        if (k == 0)
            box_spatial[long(i) * ny + j] =
                b.k0 == 0 ? spatial[long(b.i0 + i) * fny + (b.j0 + j)] : 0.0;
    }
}

__global__ void __launch_bounds__(THREADS)
scatter_dose_kernel(double* __restrict__ dose, int fny, int fnz, CellBox b,
                    double* __restrict__ local)
{
    const int nx = b.i1 - b.i0, ny = b.j1 - b.j0, nz = b.k1 - b.k0;
    const long n = long(nx) * ny * nz;
    for (long l = long(blockIdx.x) * THREADS + threadIdx.x; l < n;
         l += long(gridDim.x) * THREADS) {
        const int k = int(l % nz), j = int((l / nz) % ny), i = int(l / (long(ny) * nz));
        const long g = (long(b.i0 + i) * fny + (b.j0 + j)) * fnz + (b.k0 + k);
        dose[g] += local[l];
        local[l] = 0.0;
    }
}

__global__ void reset_kernel(BoxReduce* r)
{
    if (threadIdx.x == 0) {
        r->phi_max = 0;
        for (int a = 0; a < 3; ++a) {
            r->lo[a] = INT_MAX;
            r->hi[a] = -1;
        }
    }
}

__global__ void __launch_bounds__(THREADS)
phi_max_kernel(const double* __restrict__ phi, long n, BoxReduce* r)
{
    __shared__ unsigned long long red[THREADS / 64];
    unsigned long long k = 0;
    for (long l = long(blockIdx.x) * THREADS + threadIdx.x; l < n;
         l += long(gridDim.x) * THREADS) {
        const unsigned long long q = order_key(phi[l]);
        k = q > k ? q : k;
    }
    k = wave_max(k);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = 0;
        for (int w = 0; w < THREADS / 64; ++w)
            b = red[w] > b ? red[w] : b;
        atomicMax(&r->phi_max, b);
    }
}

// Python's max(a, b): b only when b > a.
__device__ inline double py_max(double a, double b) { return b > a ? b : a; }

__global__ void __launch_bounds__(THREADS)
significant_kernel(const double* __restrict__ phi, GridDims g, const double* __restrict__ spatial,
                   double ed, int inflow_on, double phi_in_max, double tol, BoxReduce* r)
{
    __shared__ int red[6][THREADS / 64];
    double scale = key_value(r->phi_max);
    if (inflow_on)
        scale = py_max(scale, phi_in_max);
    const double thr = tol * scale;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long n = long(nx) * ny * nz;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {-1, -1, -1};
    for (long l = long(blockIdx.x) * THREADS + threadIdx.x; l < n;
         l += long(gridDim.x) * THREADS) {
        const int k = int(l % nz), j = int((l / nz) % ny), i = int(l / (long(ny) * nz));
        bool act = phi[l] > thr;
        if (inflow_on && k == 0)
            act = act || spatial[long(i) * ny + j] * ed > thr;
        if (act) {
            lo[0] = min(lo[0], i);
            lo[1] = min(lo[1], j);
            lo[2] = min(lo[2], k);
            hi[0] = max(hi[0], i);
            hi[1] = max(hi[1], j);
            hi[2] = max(hi[2], k);
        }
    }
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min_i(lo[a]);
        hi[a] = wave_max_i(hi[a]);
    }
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) {
            red[a][threadIdx.x >> 6] = lo[a];
            red[3 + a][threadIdx.x >> 6] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        int v = a < 3 ? INT_MAX : -1;
        for (int w = 0; w < THREADS / 64; ++w)
            v = a < 3 ? min(v, red[a][w]) : max(v, red[a][w]);
        if (a < 3) {
            if (v != INT_MAX)
                atomicMin(&r->lo[a], v);
        } else if (v >= 0) {
            atomicMax(&r->hi[a - 3], v);
        }
    }
}

__global__ void __launch_bounds__(THREADS)
activity_kernel(const double* __restrict__ phi, GridDims g, const double* __restrict__ spatial,
                double ed, double phi_in_max, double tol, const BoxReduce* r,
                uint8_t* __restrict__ act)
{
    const double scale = py_max(key_value(r->phi_max), phi_in_max);
    const double thr = tol * scale;
    const int ny = g.ny, nz = g.nz;
    const long n = long(g.nx) * ny * nz;
    for (long l = long(blockIdx.x) * THREADS + threadIdx.x; l < n;
         l += long(gridDim.x) * THREADS) {
        bool a = true;
        if (scale != 0.0) {
            a = phi[l] > thr;
            if (l % nz == 0)
                a = a || spatial[l / nz] * ed > thr;
        }
        act[l] = a ? 1 : 0;
    }
}

long cells(CellBox b) { return long(b.i1 - b.i0) * (b.j1 - b.j0) * (b.k1 - b.k0); }

}  // namespace

hipError_t launch_copy(const double* src, CellBox sb, double* dst, CellBox db, hipStream_t s)
{
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(4 * cells(db))), dim3(THREADS), 0, s, src, sb,
                       dst, db);
    return hipGetLastError();
}

hipError_t launch_gather(const uint8_t* ids, const double* spatial, int ny, int nz, CellBox b,
                         uint8_t* box_ids, double* box_spatial, hipStream_t s)
{
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(cells(b))), dim3(THREADS), 0, s, ids, spatial,
                       ny, nz, b, box_ids, box_spatial);
    return hipGetLastError();
}

hipError_t launch_scatter_dose(double* dose, int ny, int nz, CellBox b, double* local,
                               hipStream_t s)
{
    hipLaunchKernelGGL(scatter_dose_kernel, dim3(grid_for(cells(b))), dim3(THREADS), 0, s, dose,
                       ny, nz, b, local);
    return hipGetLastError();
}

hipError_t launch_reset(BoxReduce* r, hipStream_t s)
{
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, r);
    return hipGetLastError();
}

hipError_t launch_phi_max(const double* phi, long n, BoxReduce* r, hipStream_t s)
{
    hipLaunchKernelGGL(phi_max_kernel, dim3(grid_for(n)), dim3(THREADS), 0, s, phi, n, r);
    return hipGetLastError();
}

hipError_t launch_significant(const double* phi, GridDims g, const double* spatial, double ed,
                              int inflow_on, double phi_in_max, double tol, BoxReduce* r,
                              hipStream_t s)
{
    const long n = long(g.nx) * g.ny * g.nz;
    hipLaunchKernelGGL(significant_kernel, dim3(grid_for(n)), dim3(THREADS), 0, s, phi, g,
                       spatial, ed, inflow_on, phi_in_max, tol, r);
    return hipGetLastError();
}

hipError_t launch_activity(const double* phi, GridDims g, const double* spatial, double ed,
                           double phi_in_max, double tol, const BoxReduce* r, uint8_t* act,
                           hipStream_t s)
{
    const long n = long(g.nx) * g.ny * g.nz;
    hipLaunchKernelGGL(activity_kernel, dim3(grid_for(n)), dim3(THREADS), 0, s, phi, g, spatial,
                       ed, phi_in_max, tol, r, act);
    return hipGetLastError();
}

}  // namespace box
}  // namespace m1
