// Global 3-D gamma index on the device (no kernel counterpart in the
// reference, which calls pymedphys from metrics.py:542-606).  It evaluates
// the exhaustive sub-voxel search of proton_solver.metrics.gamma_pass_rate,
// one thread per evaluated reference voxel, with the same double arithmetic
// in the same order (compiled with -ffp-contract=off), so every voxel's
// gamma^2 -- and the pass count -- equals the host computation bit for bit.
//
// Per evaluated voxel p (reference dose >= the cutoff): for the search
// offsets o_k sorted by length r_k (in voxels: o_k / spacing), while
// (r_k / distance)^2 < gamma^2, the evaluated dose is trilinearly
// interpolated at p + o_k and gamma^2 = min(gamma^2, ((D - D_ref) / dd)^2 +
// (r_k / distance)^2); search points outside the grid are skipped (pymedphys
// fills the evaluated dose with +inf there).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "m1march.h"

namespace {

constexpr int THREADS = 256;

thread_local std::string g_gamma_error;  // per calling thread

__global__ void __launch_bounds__(THREADS)
gamma_kernel(const double* __restrict__ dose, const double* __restrict__ ref, int nx, int ny,
             int nz, const int32_t* __restrict__ pts, int n_pts,
             const double* __restrict__ offs, const double* __restrict__ dist2, int n_offs,
             double dd, unsigned long long* n_pass, unsigned long long* n_bad)
{
    const int t = blockIdx.x * THREADS + threadIdx.x;
    bool pass = false;
    bool bad = false;
    int px = 0, py = 0, pz = 0;
    if (t < n_pts) {
        px = pts[3 * t];
        py = pts[3 * t + 1];
        pz = pts[3 * t + 2];
        // points the host did not check (device-resident lists) are
        // counted as bad and never used as an index
        bad = px < 0 || px >= nx || py < 0 || py >= ny || pz < 0 || pz >= nz;
    }
    if (t < n_pts && !bad) {
        const double rv = ref[(long(px) * ny + py) * nz + pz];
        const double hi[3] = {nx - 1.0, ny - 1.0, nz - 1.0};
        const int top[3] = {nx - 2, ny - 2, nz - 2};
        const int last[3] = {nx - 1, ny - 1, nz - 1};
        const int base[3] = {px, py, pz};
        double g2 = __longlong_as_double(0x7ff0000000000000LL);  // +inf
        for (int k = 0; k < n_offs; ++k) {
            const double dt = dist2[k];
            if (!(g2 > dt))
                break;  // offsets are sorted by length: no later one can lower gamma
            int i0[3];
            double f[3];
            bool inside = true;
            for (int a = 0; a < 3; ++a) {
                double p = double(base[a]) + offs[3 * k + a];
                inside = inside && p >= 0.0 && p <= hi[a];
                p = p < 0.0 ? 0.0 : p;
                p = p > hi[a] ? hi[a] : p;
                int i = int(floor(p));
                i = i < top[a] ? i : top[a];
                i = i > 0 ? i : 0;
                i0[a] = i;
                f[a] = p - double(i);
            }
            double val = 0.0;
            for (int dx = 0; dx < 2; ++dx) {
                const double wx = dx ? f[0] : 1.0 - f[0];
                const int ix = i0[0] + dx < last[0] ? i0[0] + dx : last[0];
                for (int dy = 0; dy < 2; ++dy) {
                    const double wy = dy ? f[1] : 1.0 - f[1];
                    const int iy = i0[1] + dy < last[1] ? i0[1] + dy : last[1];
                    for (int dz = 0; dz < 2; ++dz) {
                        const double wz = dz ? f[2] : 1.0 - f[2];
                        const int iz = i0[2] + dz < last[2] ? i0[2] + dz : last[2];
                        val += wx * wy * wz * dose[(long(ix) * ny + iy) * nz + iz];
                    }
                }
            }
            const double q = (val - rv) / dd;
            const double g = q * q + dt;
            if (inside)  // a search point outside the grid never matches
                g2 = g < g2 ? g : g2;
        }
        pass = g2 <= 1.0;
    }
    // one atomic per wave
    const unsigned long long ballot = __ballot(pass);
    if ((threadIdx.x & 63) == 0 && ballot)
        atomicAdd(n_pass, (unsigned long long)__popcll(ballot));
    const unsigned long long bad_ballot = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && bad_ballot)
        atomicAdd(n_bad, (unsigned long long)__popcll(bad_ballot));
}

// ---------------------------------------------------------------------
// Scratch of m1_gamma_pass_count_noalloc: the search offsets and their
// distance terms (staged from the host) and the two counters, in one
// file-scope device array per GPU.  Host code never passes `scratch`
// itself (its host-side name is not a device address): it passes the
// address hipGetSymbolAddress returns for the current device.
constexpr int NOALLOC_MAX_OFFSETS = M1_GAMMA_NOALLOC_MAX_OFFSETS;
constexpr size_t SCRATCH_OFFS = 0;                                       // 3 x max doubles
constexpr size_t SCRATCH_DIST = SCRATCH_OFFS + 3 * sizeof(double) * NOALLOC_MAX_OFFSETS;
constexpr size_t SCRATCH_COUNT = SCRATCH_DIST + sizeof(double) * NOALLOC_MAX_OFFSETS;
constexpr size_t SCRATCH_BYTES = SCRATCH_COUNT + 256;                    // n_pass, n_bad
static_assert(SCRATCH_DIST % 256 == 0 && SCRATCH_COUNT % 256 == 0, "aligned sections");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

__global__ void zero_counters_kernel(unsigned long long* counters)
{
    if (threadIdx.x < 2)
        counters[threadIdx.x] = 0ull;
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `device` (made current by the caller),
// looked up once per device.
hipError_t scratch_address(int device, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

template <class T>
hipError_t upload(T** d, const T* h, size_t n)
{
    hipError_t e = hipMalloc(reinterpret_cast<void**>(d), n * sizeof(T) + 16);
    if (e != hipSuccess)
        return e;
    return hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

extern "C" {

const char* m1_gamma_last_error(void) { return g_gamma_error.c_str(); }

int m1_gamma_pass_count(int32_t device, const double* dose, const double* reference, int32_t nx,
                        int32_t ny, int32_t nz, const int32_t* points, int32_t n_points,
                        const double* offsets, const double* dist_terms, int32_t n_offsets,
                        double dose_tolerance, int64_t* n_pass)
{
    g_gamma_error.clear();
    if (!dose || !reference || !n_pass || (n_points > 0 && !points) ||
        (n_offsets > 0 && (!offsets || !dist_terms))) {
        g_gamma_error = "NULL argument";
        return M1_ERR_VALUE;
    }
    if (nx <= 0 || ny <= 0 || nz <= 0 || n_points < 0 || n_offsets < 0 ||
        !(dose_tolerance > 0.0)) {
        g_gamma_error = "invalid sizes or dose tolerance";
        return M1_ERR_VALUE;
    }
    for (int32_t i = 0; i < n_points; ++i)
        if (points[3 * i] < 0 || points[3 * i] >= nx || points[3 * i + 1] < 0 ||
            points[3 * i + 1] >= ny || points[3 * i + 2] < 0 || points[3 * i + 2] >= nz) {
            g_gamma_error = "point outside the grid";
            return M1_ERR_VALUE;
        }
    for (int32_t k = 1; k < n_offsets; ++k)
        if (dist_terms[k] < dist_terms[k - 1]) {
            g_gamma_error = "offsets must be sorted by length";
            return M1_ERR_VALUE;
        }
    *n_pass = 0;
    if (n_points == 0)
        return M1_OK;
    const size_t n = size_t(nx) * ny * nz;
    double *d_dose = nullptr, *d_ref = nullptr, *d_offs = nullptr, *d_dist = nullptr;
    int32_t* d_pts = nullptr;
    unsigned long long* d_pass = nullptr;  // n_pass, n_bad
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = upload(&d_dose, dose, n);
    if (e == hipSuccess)
        e = upload(&d_ref, reference, n);
    if (e == hipSuccess)
        e = upload(&d_pts, points, 3 * size_t(n_points));
    if (e == hipSuccess)
        e = upload(&d_offs, offsets, 3 * size_t(n_offsets));
    if (e == hipSuccess)
        e = upload(&d_dist, dist_terms, size_t(n_offsets));
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void**>(&d_pass), 2 * sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipMemset(d_pass, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) {
        const unsigned blocks = unsigned((n_points + THREADS - 1) / THREADS);
        hipLaunchKernelGGL(gamma_kernel, dim3(blocks), dim3(THREADS), 0, nullptr, d_dose, d_ref,
                           nx, ny, nz, d_pts, n_points, d_offs, d_dist, n_offsets,
                           dose_tolerance, d_pass, d_pass + 1);
        e = hipGetLastError();
    }
    unsigned long long count = 0;
    if (e == hipSuccess)
        e = hipMemcpy(&count, d_pass, sizeof count, hipMemcpyDeviceToHost);
    for (void* p : {(void*)d_dose, (void*)d_ref, (void*)d_pts, (void*)d_offs, (void*)d_dist,
                    (void*)d_pass})
        if (p)
            (void)hipFree(p);
    if (e != hipSuccess) {
        g_gamma_error = std::string("gamma: ") + hipGetErrorString(e);
        return M1_ERR_HIP;
    }
    *n_pass = int64_t(count);
    return M1_OK;
}

int m1_gamma_pass_count_noalloc(int32_t device, const double* dose, const double* reference,
                                int32_t nx, int32_t ny, int32_t nz, const int32_t* points,
                                int32_t n_points, const double* offsets,
                                const double* dist_terms, int32_t n_offsets,
                                double dose_tolerance, void* stream, int64_t* n_pass)
{
    g_gamma_error.clear();
    if (!dose || !reference || !n_pass || (n_points > 0 && !points) ||
        (n_offsets > 0 && (!offsets || !dist_terms))) {
        g_gamma_error = "NULL argument";
        return M1_ERR_VALUE;
    }
    if (device < 0 || device >= MAX_DEVICES || nx <= 0 || ny <= 0 || nz <= 0 || n_points < 0 ||
        n_offsets < 0 || !(dose_tolerance > 0.0)) {
        g_gamma_error = "invalid device, sizes or dose tolerance";
        return M1_ERR_VALUE;
    }
    if (n_offsets > NOALLOC_MAX_OFFSETS) {
        g_gamma_error = "n_offsets " + std::to_string(n_offsets) + " exceeds the " +
                        std::to_string(NOALLOC_MAX_OFFSETS) +
                        " search offsets the no-allocation scratch holds";
        return M1_ERR_VALUE;
    }
    for (int32_t k = 1; k < n_offsets; ++k)
        if (dist_terms[k] < dist_terms[k - 1]) {
            g_gamma_error = "offsets must be sorted by length";
            return M1_ERR_VALUE;
        }
    *n_pass = 0;
    if (n_points == 0)
        return M1_OK;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned char* base = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = scratch_address(device, &base);
    double* d_offs = reinterpret_cast<double*>(base + SCRATCH_OFFS);
    double* d_dist = reinterpret_cast<double*>(base + SCRATCH_DIST);
    unsigned long long* d_count = reinterpret_cast<unsigned long long*>(base + SCRATCH_COUNT);
    if (e == hipSuccess && n_offsets > 0)
        e = hipMemcpyAsync(d_offs, offsets, 3 * sizeof(double) * size_t(n_offsets),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess && n_offsets > 0)
        e = hipMemcpyAsync(d_dist, dist_terms, sizeof(double) * size_t(n_offsets),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, d_count);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const unsigned blocks = unsigned((n_points + THREADS - 1) / THREADS);
        hipLaunchKernelGGL(gamma_kernel, dim3(blocks), dim3(THREADS), 0, s, dose, reference,
                           nx, ny, nz, points, n_points, d_offs, d_dist, n_offsets,
                           dose_tolerance, d_count, d_count + 1);
        e = hipGetLastError();
    }
    unsigned long long counts[2] = {0, 0};
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts, d_count, sizeof counts, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        g_gamma_error = std::string("gamma (no allocation): ") + hipGetErrorString(e);
        return M1_ERR_HIP;
    }
    if (counts[1] != 0) {
        g_gamma_error = std::to_string(counts[1]) + " point(s) outside the grid";
        return M1_ERR_VALUE;
    }
    *n_pass = int64_t(counts[0]);
    return M1_OK;
}

}  // extern "C"
