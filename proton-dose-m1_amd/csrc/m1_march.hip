// Host side of libm1march.so: the march driver of energy_march
// (numerics/energy.py:1497-1644) and the C ABI declared in include/m1march.h.
//
// The driver keeps the reference's control flow: per step one CFL reduction on
// the device, one 8-byte read-back of the inverse step, the step-local physics
// scalars evaluated on the host in double (as the reference does in Python),
// then the boundary-trace kernel and the three fused stage kernels on one HIP
// stream.  State, stage variables and dose stay resident in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "m1_host.h"

using m1::BeamParams;
using m1::HipError;
using m1::MaterialData;
using m1::NonFinite;
using m1::NotImplemented;
using m1::Num;
using m1::StepCoef;
using m1::Timed;
using m1::ValueError;
using m1::dev_alloc;
using m1::has_sources;
using m1::hip_check;
using m1::read_numerics;
using m1::K_STAGE0;
using m1::K_STAGE1;
using m1::K_STAGE2;
using m1::K_RHS;
using m1::K_FE;
using m1::K_CFL;
using m1::K_BOUNDARY;
using m1::K_SCALE;
using m1::K_TILE_MASK;
using m1::K_TILE_ZERO;
using m1::K_COUNT;

namespace {
std::string g_create_error;
}  // namespace

struct m1_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    bool have_problem = false;
    int nx = 0, ny = 0, nz = 0;
    long N = 0;
    double dx = 0, dy = 0, dz = 0;
    std::vector<MaterialData> mats;
    double beam_energy = 0, energy_spread = 0, angular_spread = 0;

    double *U = nullptr, *V1 = nullptr, *V2 = nullptr, *work = nullptr, *dose = nullptr,
           *edep = nullptr, *bnd = nullptr, *spatial = nullptr;
    uint8_t* ids = nullptr;
    // Split form of the FP64 stages (stencil input and face fluxes, sized for
    // the whole grid), allocated on first use; split_stages selects it.
    double *win = nullptr, *fx = nullptr, *fy = nullptr, *fz = nullptr;
    bool split_stages = true;
    std::vector<double> h_spatial;  // inflow profile (nx, ny), host copy
    // The cells the FP64 kernels march: the whole grid, or the current box
    // of the moving active domain (its ids and inflow profile cut out).
    struct Domain {
        int nx = 0, ny = 0, nz = 0;
        long N = 0;
        const uint8_t* ids = nullptr;
        const double* spatial = nullptr;
    } dom;
    // moving active domain buffers, allocated on first use
    uint8_t *box_ids = nullptr, *box_act = nullptr;
    double *box_spatial = nullptr, *box_dose = nullptr;
    m1::BoxReduce* d_boxred = nullptr;
    m1::BoxReduce* h_boxred = nullptr;
    // FP32-bulk buffers, allocated on first use.
    float *U32 = nullptr, *V1_32 = nullptr, *V2_32 = nullptr, *work32 = nullptr,
          *dose_hi = nullptr, *dose_lo = nullptr, *edep32 = nullptr, *bnd32 = nullptr;
    double alpha = 1.0;
    // tile skipping (FP32 bulk)
    uint8_t* d_compute[2] = {nullptr, nullptr};
    float* d_tile_max = nullptr;
    float* d_tile_face = nullptr;
    int* d_nactive = nullptr;
    int ntx = 0, nty = 0, ntz = 0;
    int cur = 0;
    bool skip_all_computed = true;  // first step of a march evaluates every tile
    int64_t tiles_evaluated = 0, tiles_full = 0;
    m1::StepCoef32* d_step32 = nullptr;
    m1::StepCoef32* h_step32 = nullptr;
    StepCoef* d_step = nullptr;
    StepCoef* d_cfl = nullptr;
    unsigned long long* d_red = nullptr;
    StepCoef* h_step = nullptr;
    StepCoef* h_cfl = nullptr;
    int* h_nactive = nullptr;
    int pending_tiles = 0;
    unsigned long long* h_red = nullptr;

    m1::Profiler prof;

    void free_problem();
    ~m1_ctx();
};

void m1_ctx::free_problem()
{
    for (void* p : {(void*)U, (void*)V1, (void*)V2, (void*)work, (void*)dose, (void*)edep,
                    (void*)bnd, (void*)spatial, (void*)ids, (void*)U32, (void*)V1_32,
                    (void*)V2_32, (void*)work32, (void*)dose_hi, (void*)dose_lo, (void*)edep32,
                    (void*)bnd32, (void*)d_compute[0], (void*)d_compute[1], (void*)d_tile_max,
                    (void*)d_tile_face, (void*)d_nactive, (void*)box_ids, (void*)box_act,
                    (void*)box_spatial, (void*)box_dose, (void*)win, (void*)fx,
                    (void*)fy, (void*)fz})
        if (p)
            (void)hipFree(p);
    U = V1 = V2 = work = dose = edep = bnd = spatial = nullptr;
    win = fx = fy = fz = nullptr;
    U32 = V1_32 = V2_32 = work32 = dose_hi = dose_lo = edep32 = bnd32 = nullptr;
    d_compute[0] = d_compute[1] = nullptr;
    d_tile_max = d_tile_face = nullptr;
    d_nactive = nullptr;
    ids = nullptr;
    box_ids = box_act = nullptr;
    box_spatial = box_dose = nullptr;
    dom = Domain{};
    have_problem = false;
}

m1_ctx::~m1_ctx()
{
    (void)hipSetDevice(device);
    if (stream)
        (void)hipStreamSynchronize(stream);
    free_problem();
    for (void* p : {(void*)d_step, (void*)d_cfl, (void*)d_red, (void*)d_step32, (void*)d_boxred})
        if (p)
            (void)hipFree(p);
    for (void* p : {(void*)h_step, (void*)h_cfl, (void*)h_red, (void*)h_step32, (void*)h_nactive,
                    (void*)h_boxred})
        if (p)
            (void)hipHostFree(p);
    prof.release();
    if (stream)
        (void)hipStreamDestroy(stream);
}

namespace {

// Reach of the WENO-Z reconstruction (ghost width 3, energy.py:612-619): the
// dilation radius of the vacuum-aware CFL activity.
constexpr int STENCIL_REACH = 3;

void sync(m1_ctx* c)
{
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (c->prof.on)
        c->prof.collect();
}

// ------------------------------------------------------------ numerics
// Dimensions of the marched domain (the whole grid except inside
// m1_march_moving, which marches a box of it).
m1::GridDims dims(const m1_ctx* c)
{
    return {c->dom.nx, c->dom.ny, c->dom.nz, c->dx, c->dy, c->dz};
}

BeamParams beam_params(const m1_ctx* c)
{
    return {c->beam_energy, c->energy_spread, c->angular_spread};
}

// All scalars of one step at (energy, dE).
void fill_step(const m1_ctx* c, const Num& n, double energy, double dE, bool ifac, bool strang,
               StepCoef& s)
{
    m1::fill_step(c->mats, beam_params(c), n, energy, dE, ifac, strang, s);
}

hipError_t stage(const m1_ctx* c, const Num& n, int kind, bool ifac, const m1::StageArgs& a)
{
    if (n.precision == M1_FP64)
        return m1::exact::launch_stage(kind, ifac, a, c->stream);
    return m1::fast::launch_stage(kind, ifac, a, c->stream);
}

// Buffers of the split stage form (m1_kernels.hip launch_split), sized for
// the whole grid so that every box of the moving domain fits.
void ensure_split(m1_ctx* c)
{
    if (c->win)
        return;
    const size_t nx = size_t(c->nx), ny = size_t(c->ny), nz = size_t(c->nz);
    dev_alloc(&c->win, 4 * nx * ny * nz, "hipMalloc stencil input");
    dev_alloc(&c->fx, 4 * (nx + 1) * ny * nz, "hipMalloc x fluxes");
    dev_alloc(&c->fy, 4 * nx * (ny + 1) * nz, "hipMalloc y fluxes");
    dev_alloc(&c->fz, 4 * nx * ny * (nz + 1), "hipMalloc z fluxes");
}

m1::StageArgs base_args(m1_ctx* c, const Num& n)
{
    m1::StageArgs a{};
    a.g = dims(c);
    a.ids = c->dom.ids;
    a.coef = c->d_step;
    a.wave_mode = n.wave;
    if (c->split_stages) {
        ensure_split(c);
        a.win = c->win;
        a.fx = c->fx;
        a.fy = c->fy;
        a.fz = c->fz;
    }
    return a;
}

void upload_step(m1_ctx* c)
{
    hip_check(hipMemcpyAsync(c->d_step, c->h_step, sizeof(StepCoef), hipMemcpyHostToDevice,
                             c->stream),
              "coefficient upload");
}

void run_boundary(m1_ctx* c, const Num& n, bool ifac)
{
    m1::BoundaryArgs b{dims(c), c->dom.spatial, c->dom.ids, c->d_step, c->bnd, ifac ? 1 : 0};
    Timed t(c->prof, c->stream, K_BOUNDARY);
    hipError_t e = n.precision == M1_FP64 ? m1::exact::launch_boundary(b, c->stream)
                                                 : m1::fast::launch_boundary(b, c->stream);
    hip_check(e, "boundary kernel");
}

void run_scale(m1_ctx* c, const Num& n)
{
    m1::ScaleArgs s{c->dom.N, c->U, c->dom.ids, c->d_step, n.nuclear ? 1 : 0,
                    n.scattering != M1_SCATTERING_OFF ? 1 : 0};
    Timed t(c->prof, c->stream, K_SCALE);
    hip_check(m1::fast::launch_scale(s, c->stream), "source half-step kernel");
}

// Three fused SSPRK3 stages (energy.py:725-873 / 878-1130).  Requires the
// step coefficients on the device.
void rk3(m1_ctx* c, const Num& n, bool ifac, double* dose, double* edep)
{
    run_boundary(c, n, ifac);
    const long m2 = long(c->dom.nx) * c->dom.ny;
    m1::StageArgs a = base_args(c, n);
    a.U = c->U;
    a.V1 = c->V1;
    {
        a.src = c->U;
        a.out = c->V1;
        a.bnd = c->bnd;
        Timed t(c->prof, c->stream, K_STAGE0);
        hip_check(stage(c, n, m1::STAGE0, false, a), "stage 0 kernel");
    }
    {
        a.src = c->V1;
        a.out = c->V2;
        a.bnd = c->bnd + 4 * m2;
        Timed t(c->prof, c->stream, K_STAGE1);
        hip_check(stage(c, n, m1::STAGE1, ifac, a), "stage 1 kernel");
    }
    {
        a.src = c->V2;
        a.out = nullptr;
        a.dose = dose;
        a.edep = edep;
        a.bnd = c->bnd + 8 * m2;
        Timed t(c->prof, c->stream, K_STAGE2);
        hip_check(stage(c, n, m1::STAGE2, ifac, a), "stage 2 kernel");
    }
}

// forward_euler (energy.py:663-722): new state into `work`, then swap.
void forward_euler(m1_ctx* c, const Num& n, double* dose, double* edep)
{
    run_boundary(c, n, false);
    m1::StageArgs a = base_args(c, n);
    a.src = c->U;
    a.U = c->U;
    a.out = c->work;
    a.dose = dose;
    a.edep = edep;
    a.bnd = c->bnd;
    {
        Timed t(c->prof, c->stream, K_FE);
        hip_check(stage(c, n, m1::STAGE_FE, false, a), "forward Euler kernel");
    }
    std::swap(c->U, c->work);
}

// advance_energy_step (energy.py:1132-1165).
void advance(m1_ctx* c, const Num& n, double* dose, double* edep)
{
    if (n.scheme == M1_FORWARD_EULER)
        forward_euler(c, n, dose, edep);
    else
        rk3(c, n, false, dose, edep);
}

// One step of the requested kind; coefficients are filled and uploaded here.
void step(m1_ctx* c, const Num& n, int kind, double energy, double dE, double* dose,
          double* edep)
{
    if (dE <= 0.0 && kind != M1_STEP_FORWARD_EULER)
        throw ValueError{"dE must be positive."};
    const bool ifac = kind == M1_STEP_IF_SSPRK3 && has_sources(n);
    const bool strang = kind == M1_STEP_STRANG;
    fill_step(c, n, energy, dE, ifac, strang, *c->h_step);
    upload_step(c);
    switch (kind) {
    case M1_STEP_FORWARD_EULER:
        forward_euler(c, n, dose, edep);
        break;
    case M1_STEP_SSPRK3:
        rk3(c, n, false, dose, edep);
        break;
    case M1_STEP_IF_SSPRK3:
        rk3(c, n, ifac, dose, edep);
        break;
    case M1_STEP_STRANG:
        if (has_sources(n))
            run_scale(c, n);
        advance(c, n, dose, edep);
        if (has_sources(n))
            run_scale(c, n);
        break;
    default:
        throw ValueError{"unknown step kind"};
    }
}

// compute_energy_step (energy.py:539-660) of the device state at `energy`;
// `act` (nullable) is the cell activity of the vacuum-aware CFL.
double cfl_step(m1_ctx* c, const Num& n, double energy, const uint8_t* act = nullptr)
{
    StepCoef& s = *c->h_cfl;
    m1::fill_cfl(c->mats, beam_params(c), energy, s);
    hip_check(hipMemcpyAsync(c->d_cfl, c->h_cfl, sizeof(StepCoef), hipMemcpyHostToDevice,
                             c->stream),
              "CFL coefficient upload");
    hip_check(hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long), c->stream),
              "CFL reset");
    m1::CflArgs a{dims(c), c->U,   c->dom.ids, c->dom.spatial, c->d_cfl,
                  c->d_red, n.wave, act,       act ? STENCIL_REACH : 0};
    {
        Timed t(c->prof, c->stream, K_CFL);
        hipError_t e = n.precision == M1_FP64 ? m1::exact::launch_cfl(a, c->stream)
                                                     : m1::fast::launch_cfl(a, c->stream);
        hip_check(e, "CFL kernel");
    }
    hip_check(hipMemcpyAsync(c->h_red, c->d_red, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, c->stream),
              "CFL read-back");
    sync(c);
    double mx;
    std::memcpy(&mx, c->h_red, sizeof mx);
    return m1::cfl_energy_step(n, mx);
}

void require_problem(const m1_ctx* c)
{
    if (!c->have_problem)
        throw ValueError{"no problem bound: call m1_set_problem first"};
}

template <class F>
int guarded(m1_ctx* c, F&& f)
{
    if (!c)
        return M1_ERR_STATE;
    return m1::run_guarded(c->device, c->err, f);
}


// ------------------------------------------------------------- FP32 bulk
// The FP32-bulk march keeps the device state as U / alpha in float, with
// alpha the peak inflow fluence max(profile) * N(E_beam) (so phi <= ~1).

void ensure_f32(m1_ctx* c)
{
    if (c->U32)
        return;
    const size_t N = size_t(c->N), m2 = size_t(c->nx) * c->ny;
    dev_alloc(&c->U32, 4 * N, "hipMalloc state (f32)");
    dev_alloc(&c->V1_32, 4 * N, "hipMalloc stage 1 (f32)");
    dev_alloc(&c->V2_32, 4 * N, "hipMalloc stage 2 (f32)");
    dev_alloc(&c->work32, 4 * N, "hipMalloc work (f32)");
    dev_alloc(&c->dose_hi, N, "hipMalloc dose (f32)");
    dev_alloc(&c->dose_lo, N, "hipMalloc dose (f32)");
    dev_alloc(&c->edep32, N, "hipMalloc deposition (f32)");
    dev_alloc(&c->bnd32, 12 * m2, "hipMalloc boundary (f32)");
    c->ntx = (c->nx + m1::F32_TX - 1) / m1::F32_TX;
    c->nty = (c->ny + m1::F32_TY - 1) / m1::F32_TY;
    c->ntz = (c->nz + m1::F32_TZ - 1) / m1::F32_TZ;
    const size_t nt = size_t(c->ntx) * c->nty * c->ntz;
    dev_alloc(&c->d_compute[0], nt, "hipMalloc tile set");
    dev_alloc(&c->d_compute[1], nt, "hipMalloc tile set");
    dev_alloc(&c->d_tile_max, nt, "hipMalloc tile max");
    dev_alloc(&c->d_tile_face, nt * m1::FACE_SLOTS, "hipMalloc tile face slabs");
    dev_alloc(&c->d_nactive, 1, "hipMalloc tile count");
    if (!c->d_step32) {
        dev_alloc(&c->d_step32, 1, "hipMalloc coefficients (f32)");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_step32), sizeof(m1::StepCoef32)),
                  "hipHostMalloc");
    }
}

void upload_step32(m1_ctx* c)
{
    m1::set_norm(c->alpha, *c->h_step);
    m1::fill_step32(c->mats.size(), c->alpha, *c->h_step, *c->h_step32);
    hip_check(hipMemcpyAsync(c->d_step, c->h_step, sizeof(StepCoef), hipMemcpyHostToDevice,
                             c->stream),
              "coefficient upload");
    hip_check(hipMemcpyAsync(c->d_step32, c->h_step32, sizeof(m1::StepCoef32),
                             hipMemcpyHostToDevice, c->stream),
              "coefficient upload (f32)");
}

m1::StageArgs32 base_args32(const m1_ctx* c, const Num& n)
{
    m1::StageArgs32 a{};
    a.g = dims(c);
    a.ids = c->ids;
    a.coef = c->d_step32;
    a.wave_mode = n.wave;
    a.f64_closure = n.precision == M1_FP32_BULK_F64_CLOSURE ? 1 : 0;
    return a;
}

// Tile set of this step (nullptr: evaluate every tile).
const uint8_t* step_tiles(const m1_ctx* c, const Num& n)
{
    if (n.tile_threshold <= 0.0 || c->skip_all_computed)
        return nullptr;
    return c->d_compute[c->cur];
}

void rk3_f32(m1_ctx* c, const Num& n, bool ifac, bool dose, bool edep)
{
    {
        m1::BoundaryArgs32 b{dims(c), c->spatial, c->ids, c->d_step, c->bnd32, ifac ? 1 : 0};
        Timed t(c->prof, c->stream, K_BOUNDARY);
        hip_check(m1::f32::launch_boundary(b, c->stream), "boundary kernel (f32)");
    }
    const long m2 = long(c->nx) * c->ny;
    m1::StageArgs32 a = base_args32(c, n);
    a.U = c->U32;
    a.V1 = c->V1_32;
    a.compute = step_tiles(c, n);
    a.tile_max = n.tile_threshold > 0.0 ? c->d_tile_max : nullptr;
    a.tile_face = n.tile_threshold > 0.0 ? c->d_tile_face : nullptr;
    {
        a.src = c->U32;
        a.out = c->V1_32;
        a.bnd = c->bnd32;
        Timed t(c->prof, c->stream, K_STAGE0);
        hip_check(m1::f32::launch_stage(m1::STAGE0, false, a, c->stream), "stage 0 kernel (f32)");
    }
    {
        a.src = c->V1_32;
        a.out = c->V2_32;
        a.bnd = c->bnd32 + 4 * m2;
        Timed t(c->prof, c->stream, K_STAGE1);
        hip_check(m1::f32::launch_stage(m1::STAGE1, ifac, a, c->stream), "stage 1 kernel (f32)");
    }
    {
        a.src = c->V2_32;
        a.out = nullptr;
        a.dose_hi = dose ? c->dose_hi : nullptr;
        a.dose_lo = dose ? c->dose_lo : nullptr;
        a.edep = edep ? c->edep32 : nullptr;
        a.bnd = c->bnd32 + 8 * m2;
        Timed t(c->prof, c->stream, K_STAGE2);
        hip_check(m1::f32::launch_stage(m1::STAGE2, ifac, a, c->stream), "stage 2 kernel (f32)");
    }
}

void forward_euler_f32(m1_ctx* c, const Num& n, bool dose, bool edep)
{
    {
        m1::BoundaryArgs32 b{dims(c), c->spatial, c->ids, c->d_step, c->bnd32, 0};
        Timed t(c->prof, c->stream, K_BOUNDARY);
        hip_check(m1::f32::launch_boundary(b, c->stream), "boundary kernel (f32)");
    }
    m1::StageArgs32 a = base_args32(c, n);
    a.src = c->U32;
    a.U = c->U32;
    a.out = c->work32;
    a.dose_hi = dose ? c->dose_hi : nullptr;
    a.dose_lo = dose ? c->dose_lo : nullptr;
    a.edep = edep ? c->edep32 : nullptr;
    a.bnd = c->bnd32;
    {
        Timed t(c->prof, c->stream, K_FE);
        hip_check(m1::f32::launch_stage(m1::STAGE_FE, false, a, c->stream),
                  "forward Euler kernel (f32)");
    }
    std::swap(c->U32, c->work32);
}

void run_scale_f32(m1_ctx* c, const Num& n)
{
    Timed t(c->prof, c->stream, K_SCALE);
    hip_check(m1::f32::launch_scale(c->N, c->U32, c->ids, c->d_step, n.nuclear ? 1 : 0,
                                    n.scattering != M1_SCATTERING_OFF ? 1 : 0, c->stream),
              "source half-step kernel (f32)");
}

void step_f32(m1_ctx* c, const Num& n, int kind, double energy, double dE, bool dose, bool edep)
{
    if (dE <= 0.0 && kind != M1_STEP_FORWARD_EULER)
        throw ValueError{"dE must be positive."};
    const bool ifac = kind == M1_STEP_IF_SSPRK3 && has_sources(n);
    const bool strang = kind == M1_STEP_STRANG;
    fill_step(c, n, energy, dE, ifac, strang, *c->h_step);
    upload_step32(c);
    auto adv = [&] {
        if (n.scheme == M1_FORWARD_EULER)
            forward_euler_f32(c, n, dose, edep);
        else
            rk3_f32(c, n, false, dose, edep);
    };
    switch (kind) {
    case M1_STEP_FORWARD_EULER:
        forward_euler_f32(c, n, dose, edep);
        break;
    case M1_STEP_SSPRK3:
        rk3_f32(c, n, false, dose, edep);
        break;
    case M1_STEP_IF_SSPRK3:
        rk3_f32(c, n, ifac, dose, edep);
        break;
    case M1_STEP_STRANG:
        if (has_sources(n))
            run_scale_f32(c, n);
        adv();
        if (has_sources(n))
            run_scale_f32(c, n);
        break;
    default:
        throw ValueError{"unknown step kind"};
    }
}

double cfl_step_f32(m1_ctx* c, const Num& n, double energy)
{
    StepCoef& s = *c->h_cfl;
    m1::fill_cfl(c->mats, beam_params(c), energy, s);
    m1::set_norm(c->alpha, s);
    hip_check(hipMemcpyAsync(c->d_cfl, c->h_cfl, sizeof(StepCoef), hipMemcpyHostToDevice,
                             c->stream),
              "CFL coefficient upload");
    hip_check(hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long), c->stream), "CFL reset");
    m1::CflArgs32 a{dims(c), c->U32, c->ids, c->spatial, c->d_cfl, c->d_red, n.wave};
    {
        Timed t(c->prof, c->stream, K_CFL);
        hip_check(m1::f32::launch_cfl(a, c->stream), "CFL kernel (f32)");
    }
    hip_check(hipMemcpyAsync(c->h_red, c->d_red, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, c->stream),
              "CFL read-back");
    sync(c);
    if (c->pending_tiles) {
        c->tiles_evaluated += *c->h_nactive;
        c->pending_tiles = 0;
    }
    double mx;
    std::memcpy(&mx, c->h_red, sizeof mx);
    return m1::cfl_energy_step(n, mx);
}

// Next step's tile set from this step's tile maxima (rk3 steps only; the
// forward-Euler and Strang paths evaluate every tile).
void update_tiles(m1_ctx* c, const Num& n, double energy_next)
{
    const int64_t nt = int64_t(c->ntx) * c->nty * c->ntz;
    c->tiles_full += nt;
    c->tiles_evaluated += c->skip_all_computed ? nt : 0;
    if (n.tile_threshold <= 0.0)
        return;
    if (c->skip_all_computed) {
        // the first step evaluated every tile: mark all as computed
        hip_check(hipMemsetAsync(c->d_compute[c->cur], 1, size_t(nt), c->stream), "tile set");
    }
    // Inflow spectrum relative to its peak anywhere in the next two maximal steps.
    const double lo = energy_next - 2.0 * n.max_step;
    const double e = std::min(std::max(c->beam_energy, lo), energy_next);
    const double d = (e - c->beam_energy) / c->energy_spread;
    const int inflow_on = std::exp(-0.5 * d * d) > n.tile_threshold ? 1 : 0;
    hip_check(hipMemsetAsync(c->d_nactive, 0, sizeof(int), c->stream), "tile count reset");
    m1::TileMaskArgs ma{c->ntx, c->nty, c->ntz, c->d_tile_max, c->d_tile_face, c->d_compute[c->cur],
                        c->d_compute[1 - c->cur], float(n.tile_threshold), inflow_on,
                        m1::tile_radius(m1::F32_TX), m1::tile_radius(m1::F32_TY),
                        m1::tile_radius(m1::F32_TZ), c->d_nactive};
    {
        Timed t(c->prof, c->stream, K_TILE_MASK);
        hip_check(m1::f32::launch_tile_mask(ma, c->stream), "tile mask kernel");
    }
    m1::TileZeroArgs za{dims(c), c->d_compute[c->cur], c->d_compute[1 - c->cur], c->U32,
                        c->V1_32, c->V2_32};
    {
        Timed t(c->prof, c->stream, K_TILE_ZERO);
        hip_check(m1::f32::launch_tile_zero(za, c->stream), "tile zero kernel");
    }
    hip_check(hipMemcpyAsync(c->h_nactive, c->d_nactive, sizeof(int), hipMemcpyDeviceToHost,
                             c->stream),
              "tile count read-back");
    c->pending_tiles = 1;
    c->cur = 1 - c->cur;
    c->skip_all_computed = false;
}

// Move the FP64 device state into the scaled float state and back.
void narrow_state(m1_ctx* c)
{
    ensure_f32(c);
    c->skip_all_computed = true;
    c->cur = 0;
    c->pending_tiles = 0;
    hip_check(m1::f32::launch_narrow_state(c->U, 1.0 / c->alpha, c->U32, 4 * c->N, c->stream),
              "state narrowing");
}

void widen_state(m1_ctx* c, double* out)
{
    hip_check(m1::f32::launch_widen_state(c->U32, c->alpha, out, 4 * c->N, c->stream),
              "state widening");
}

// ------------------------------------------------- moving active domain
// _moving_active_energy_march (energy.py:1262-1494) with the box helpers of
// moving_active_domain.py:35-253.  The state lives on a box of the grid that
// follows the significant cells; the box is made the marched domain (its ids
// and inflow profile cut out by a kernel), so the step kernels run on it
// unchanged, and the box's dose is added into the full-grid dose after every
// step.
using m1::CellBox;

CellBox full_box(const m1_ctx* c) { return {0, c->nx, 0, c->ny, 0, c->nz}; }

long cells(const CellBox& b) { return long(b.i1 - b.i0) * (b.j1 - b.j0) * (b.k1 - b.k0); }

bool same_box(const CellBox& a, const CellBox& b)
{
    return a.i0 == b.i0 && a.i1 == b.i1 && a.j0 == b.j0 && a.j1 == b.j1 && a.k0 == b.k0 &&
           a.k1 == b.k1;
}

// initial_box (moving_active_domain.py:35-60): the beam's transverse support
// widened by `guard`, z in [0, 1 + guard).
CellBox initial_box(const m1_ctx* c, int guard)
{
    int ix0 = -1, ix1 = -1, iy0 = c->ny, iy1 = -1;
    for (int i = 0; i < c->nx; ++i)
        for (int j = 0; j < c->ny; ++j)
            if (c->h_spatial[size_t(i) * c->ny + j] > 0.0) {
                if (ix0 < 0)
                    ix0 = i;
                ix1 = i;
                iy0 = std::min(iy0, j);
                iy1 = std::max(iy1, j);
            }
    if (ix0 < 0)
        throw ValueError{"Beam inflow has empty transverse support."};
    return {std::max(0, ix0 - guard), std::min(c->nx, ix1 + 1 + guard),
            std::max(0, iy0 - guard), std::min(c->ny, iy1 + 1 + guard),
            0,                        std::min(c->nz, 1 + guard)};
}

// Largest inflow phi of box b at spectrum value ed: max over the box's
// columns of profile * ed (boundary.py:82-114), zero once the box has left
// z = 0 (make_local_problem, moving_active_domain.py:84-92).
double inflow_phi_max(const m1_ctx* c, const CellBox& b, double ed)
{
    if (b.k0 != 0)
        return 0.0;
    double mx = c->h_spatial[size_t(b.i0) * c->ny + b.j0] * ed;
    for (int i = b.i0; i < b.i1; ++i)
        for (int j = b.j0; j < b.j1; ++j) {
            const double v = c->h_spatial[size_t(i) * c->ny + j] * ed;
            if (v > mx)
                mx = v;
        }
    return mx;
}

void ensure_box_buffers(m1_ctx* c)
{
    if (c->box_ids)
        return;
    dev_alloc(&c->box_ids, size_t(c->N), "hipMalloc box ids");
    dev_alloc(&c->box_act, size_t(c->N), "hipMalloc box activity");
    dev_alloc(&c->box_spatial, size_t(c->nx) * c->ny, "hipMalloc box inflow");
    dev_alloc(&c->box_dose, size_t(c->N), "hipMalloc box dose");
    if (!c->d_boxred)
        dev_alloc(&c->d_boxred, 1, "hipMalloc box reduction");
    if (!c->h_boxred)
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_boxred), sizeof(m1::BoxReduce)),
                  "hipHostMalloc");
    hip_check(hipMemsetAsync(c->box_dose, 0, size_t(c->N) * sizeof(double), c->stream),
              "box dose reset");
}

// Make box b the marched domain.
void set_box_domain(m1_ctx* c, const CellBox& b)
{
    hip_check(m1::box::launch_gather(c->ids, c->spatial, c->ny, c->nz, b, c->box_ids,
                                     c->box_spatial, c->stream),
              "box gather kernel");
    c->dom = {b.i1 - b.i0, b.j1 - b.j0, b.k1 - b.k0, cells(b), c->box_ids, c->box_spatial};
}

void set_full_domain(m1_ctx* c) { c->dom = {c->nx, c->ny, c->nz, c->N, c->ids, c->spatial}; }

// The state U on box `from` becomes the state on box `to` (zero outside the
// overlap; moving_active_domain.py:196-244).
void move_state(m1_ctx* c, const CellBox& from, const CellBox& to)
{
    hip_check(m1::box::launch_copy(c->U, from, c->work, to, c->stream), "box copy kernel");
    std::swap(c->U, c->work);
}

// desired_box (moving_active_domain.py:154-193) of the state on box b at
// `energy`: significant_box (107-151) widened by `guard`, or initial_box
// when no cell is significant.
CellBox desired_box(m1_ctx* c, const CellBox& b, int guard, double tol, double ed)
{
    const m1::GridDims g = dims(c);
    const int inflow_on = b.k0 == 0;
    const double phi_in_max = inflow_phi_max(c, b, ed);
    hip_check(m1::box::launch_reset(c->d_boxred, c->stream), "box reduction reset");
    hip_check(m1::box::launch_phi_max(c->U, c->dom.N, c->d_boxred, c->stream), "phi max kernel");
    hip_check(m1::box::launch_significant(c->U, g, c->dom.spatial, ed, inflow_on, phi_in_max, tol,
                                          c->d_boxred, c->stream),
              "significant box kernel");
    hip_check(hipMemcpyAsync(c->h_boxred, c->d_boxred, sizeof(m1::BoxReduce),
                             hipMemcpyDeviceToHost, c->stream),
              "box reduction read-back");
    sync(c);
    const m1::BoxReduce& r = *c->h_boxred;
    double scale = m1::key_value(r.phi_max);
    if (inflow_on && phi_in_max > scale)
        scale = phi_in_max;
    if (scale == 0.0 || r.hi[0] < 0)
        return initial_box(c, guard);
    return {std::max(0, b.i0 + r.lo[0] - guard), std::min(c->nx, b.i0 + r.hi[0] + 1 + guard),
            std::max(0, b.j0 + r.lo[1] - guard), std::min(c->ny, b.j0 + r.hi[1] + 1 + guard),
            std::max(0, b.k0 + r.lo[2] - guard), std::min(c->nz, b.k0 + r.hi[2] + 1 + guard)};
}

// compute_energy_step with vacuum_relative_tolerance = tol (energy.py:493-660)
// of the state on the marched domain; `b` is that domain's box of the grid.
double cfl_step_vacuum(m1_ctx* c, const Num& n, double energy, const CellBox& b, double tol)
{
    if (n.wave != 0)
        return cfl_step(c, n, energy);  // the masks apply to the exact M1 speeds only
    const double ed = m1::energy_density(energy, c->beam_energy, c->energy_spread);
    hip_check(m1::box::launch_reset(c->d_boxred, c->stream), "box reduction reset");
    hip_check(m1::box::launch_phi_max(c->U, c->dom.N, c->d_boxred, c->stream), "phi max kernel");
    hip_check(m1::box::launch_activity(c->U, dims(c), c->dom.spatial, ed,
                                       inflow_phi_max(c, b, ed), tol, c->d_boxred, c->box_act,
                                       c->stream),
              "activity kernel");
    return cfl_step(c, n, energy, c->box_act);
}

}  // namespace

extern "C" {

int m1_abi_version(void) { return M1_ABI_VERSION; }

int m1_create(m1_ctx** out, int device)
{
    if (!out) {
        g_create_error = "out must not be NULL";
        return M1_ERR_VALUE;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_create_error = "no HIP device is available";
        return M1_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        g_create_error = "device ordinal out of range";
        return M1_ERR_VALUE;
    }
    m1_ctx* c = new m1_ctx;
    c->device = device;
    try {
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        dev_alloc(&c->d_step, 1, "hipMalloc coefficients");
        dev_alloc(&c->d_cfl, 1, "hipMalloc coefficients");
        dev_alloc(&c->d_red, 1, "hipMalloc reduction");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_step), sizeof(StepCoef)),
                  "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_cfl), sizeof(StepCoef)),
                  "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_red), sizeof(unsigned long long)),
                  "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_nactive), sizeof(int)),
                  "hipHostMalloc");
    } catch (const HipError& e) {
        g_create_error = e.msg;
        delete c;
        return M1_ERR_HIP;
    }
    *out = c;
    return M1_OK;
}

void m1_destroy(m1_ctx* ctx) { delete ctx; }

const char* m1_last_error(const m1_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int m1_set_problem(m1_ctx* ctx, int32_t nx, int32_t ny, int32_t nz, double dx, double dy,
                   double dz, const int32_t* ids, int32_t n_materials,
                   const m1_material* materials, const double* inflow_spatial,
                   const m1_beam* beam)
{
    return guarded(ctx, [&] {
        if (nx <= 0 || ny <= 0 || nz <= 0)
            throw ValueError{"Geometry resolution must contain positive integers."};
        if (nz < 4)
            throw ValueError{"WENO-Z inflow extrapolation requires at least four z cells."};
        if (!(dx > 0 && dy > 0 && dz > 0))
            throw ValueError{"cell sizes must be positive"};
        if (n_materials < 1 || n_materials > m1::MAXMAT)
            throw ValueError{"between 1 and 8 materials are supported"};
        if (!ids || !materials || !inflow_spatial || !beam)
            throw ValueError{"ids, materials, inflow and beam must not be NULL"};
        if (beam->energy_spread == 0.0)
            throw ValueError{"Zero energy spread is not supported by the Gaussian inflow."};
        const long N = long(nx) * ny * nz;
        std::vector<uint8_t> id8(static_cast<size_t>(N));
        for (long i = 0; i < N; ++i) {
            if (ids[i] < 0 || ids[i] >= n_materials)
                throw ValueError{"material id out of range"};
            id8[size_t(i)] = uint8_t(ids[i]);
        }
        std::vector<MaterialData> mats;
        for (int m = 0; m < n_materials; ++m)
            mats.push_back(m1::read_material(materials[m], m));

        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        ctx->free_problem();
        ctx->nx = nx;
        ctx->ny = ny;
        ctx->nz = nz;
        ctx->N = N;
        ctx->dx = dx;
        ctx->dy = dy;
        ctx->dz = dz;
        ctx->mats = std::move(mats);
        ctx->beam_energy = beam->energy;
        ctx->energy_spread = beam->energy_spread;
        ctx->angular_spread = beam->angular_spread;
        {
            double pk = 0.0;
            for (size_t i = 0; i < size_t(nx) * ny; ++i)
                pk = std::max(pk, inflow_spatial[i]);
            ctx->alpha = m1::march_alpha(pk, beam_params(ctx));
        }
        const size_t m2 = size_t(nx) * ny;
        dev_alloc(&ctx->U, 4 * size_t(N), "hipMalloc state");
        dev_alloc(&ctx->V1, 4 * size_t(N), "hipMalloc stage 1");
        dev_alloc(&ctx->V2, 4 * size_t(N), "hipMalloc stage 2");
        dev_alloc(&ctx->work, 4 * size_t(N), "hipMalloc work");
        dev_alloc(&ctx->dose, size_t(N), "hipMalloc dose");
        dev_alloc(&ctx->edep, size_t(N), "hipMalloc deposition");
        dev_alloc(&ctx->bnd, 12 * m2, "hipMalloc boundary");
        dev_alloc(&ctx->spatial, m2, "hipMalloc inflow");
        dev_alloc(&ctx->ids, size_t(N), "hipMalloc ids");
        hip_check(hipMemcpy(ctx->ids, id8.data(), size_t(N), hipMemcpyHostToDevice), "ids upload");
        hip_check(hipMemcpy(ctx->spatial, inflow_spatial, m2 * sizeof(double),
                            hipMemcpyHostToDevice),
                  "inflow upload");
        hip_check(hipMemsetAsync(ctx->U, 0, 4 * size_t(N) * sizeof(double), ctx->stream),
                  "state reset");
        hip_check(hipMemsetAsync(ctx->dose, 0, size_t(N) * sizeof(double), ctx->stream),
                  "dose reset");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        ctx->h_spatial.assign(inflow_spatial, inflow_spatial + m2);
        ctx->dom = {nx, ny, nz, N, ctx->ids, ctx->spatial};
        ctx->have_problem = true;
    });
}

int m1_set_state(m1_ctx* ctx, const double* U)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        const size_t bytes = 4 * size_t(ctx->N) * sizeof(double);
        if (U)
            hip_check(hipMemcpyAsync(ctx->U, U, bytes, hipMemcpyHostToDevice, ctx->stream),
                      "state upload");
        else
            hip_check(hipMemsetAsync(ctx->U, 0, bytes, ctx->stream), "state reset");
        sync(ctx);
    });
}

int m1_get_state(m1_ctx* ctx, double* U)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        if (!U)
            throw ValueError{"U must not be NULL"};
        hip_check(hipMemcpyAsync(U, ctx->U, 4 * size_t(ctx->N) * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream),
                  "state download");
        sync(ctx);
    });
}

int m1_get_dose(m1_ctx* ctx, double* dose)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        if (!dose)
            throw ValueError{"dose must not be NULL"};
        hip_check(hipMemcpyAsync(dose, ctx->dose, size_t(ctx->N) * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream),
                  "dose download");
        sync(ctx);
    });
}

int m1_zero_dose(m1_ctx* ctx)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        hip_check(hipMemsetAsync(ctx->dose, 0, size_t(ctx->N) * sizeof(double), ctx->stream),
                  "dose reset");
        sync(ctx);
    });
}

int m1_march(m1_ctx* ctx, const m1_numerics* num, double* energies, int32_t energies_cap,
             int32_t* n_energies, const double* snapshot_energies, int32_t n_snapshots,
             double* snapshots, int32_t* snapshot_taken)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        const Num n = read_numerics(num);
        if (n_snapshots < 0 || (n_snapshots > 0 && !snapshot_energies))
            throw ValueError{"invalid snapshot request"};
        double energy = ctx->beam_energy + n.window * ctx->energy_spread;
        const double emin = n.emin;
        // Unique requested energies inside [emin, E0], descending (energy.py:1533-1541).
        std::vector<double> targets;
        for (int i = 0; i < n_snapshots; ++i) {
            if (snapshot_taken)
                snapshot_taken[i] = 0;
            const double t = snapshot_energies[i];
            if (emin <= t && t <= energy &&
                std::find(targets.begin(), targets.end(), t) == targets.end())
                targets.push_back(t);
        }
        std::sort(targets.begin(), targets.end(), std::greater<double>());
        hip_check(hipMemsetAsync(ctx->dose, 0, size_t(ctx->N) * sizeof(double), ctx->stream),
                  "dose reset");
        int32_t ne = 0;
        auto record = [&](double e) {
            if (energies && ne < energies_cap)
                energies[ne] = e;
            ++ne;
        };
        record(energy);
        size_t ti = 0;
        const bool f32 = m1::is_f32(n);
        if (f32) {
            narrow_state(ctx);
            hip_check(hipMemsetAsync(ctx->dose_hi, 0, size_t(ctx->N) * sizeof(float), ctx->stream),
                      "dose reset");
            hip_check(hipMemsetAsync(ctx->dose_lo, 0, size_t(ctx->N) * sizeof(float), ctx->stream),
                      "dose reset");
        }
        const int kind = n.scheme == M1_IF_SSPRK3 ? M1_STEP_IF_SSPRK3
                         : has_sources(n)         ? M1_STEP_STRANG
                         : n.scheme == M1_FORWARD_EULER ? M1_STEP_FORWARD_EULER
                                                        : M1_STEP_SSPRK3;
        while (energy > emin) {
            double dE = f32 ? cfl_step_f32(ctx, n, energy) : cfl_step(ctx, n, energy);
            if (ti < targets.size()) {
                const double t = targets[ti];
                if (energy > t && t >= energy - dE)
                    dE = energy - t;
            }
            if (energy - emin < dE)
                dE = energy - emin;
            if (f32) {
                step_f32(ctx, n, kind, energy, dE, true, false);
                if (kind == M1_STEP_IF_SSPRK3 || kind == M1_STEP_SSPRK3)
                    update_tiles(ctx, n, energy - dE);
            } else {
                step(ctx, n, kind, energy, dE, ctx->dose, nullptr);
            }
            energy -= dE;
            record(energy);
            if (ti < targets.size()) {
                const double t = targets[ti];
                if (std::fabs(energy - t) <= 1e-12) {
                    if (snapshots) {
                        if (f32)
                            widen_state(ctx, ctx->work);
                        // slot of the first request equal to this target
                        int slot = 0;
                        for (int i = 0; i < n_snapshots; ++i)
                            if (snapshot_energies[i] == t) {
                                slot = i;
                                break;
                            }
                        hip_check(hipMemcpyAsync(snapshots + size_t(slot) * 4 * size_t(ctx->N),
                                                 f32 ? ctx->work : ctx->U,
                                                 4 * size_t(ctx->N) * sizeof(double),
                                                 hipMemcpyDeviceToHost, ctx->stream),
                                  "snapshot download");
                        if (snapshot_taken)
                            snapshot_taken[slot] = 1;
                    }
                    ++ti;
                }
            }
        }
        if (f32) {
            widen_state(ctx, ctx->U);
            hip_check(m1::f32::launch_finalize_dose(ctx->dose_hi, ctx->dose_lo, ctx->dose, ctx->N,
                                                    ctx->stream),
                      "dose finalisation");
        }
        sync(ctx);
        if (n_energies)
            *n_energies = ne;
    });
}

int m1_march_moving(m1_ctx* ctx, const m1_numerics* num, int32_t guard_cells,
                    double truncation_sigma, double* energies, int32_t energies_cap,
                    int32_t* n_energies, const double* snapshot_energies, int32_t n_snapshots,
                    double* snapshots, int32_t* snapshot_taken, int32_t* boxes)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        const Num n = read_numerics(num);
        if (m1::is_f32(n))
            throw NotImplemented{"moving_active_domain is marched in FP64 and FP64_FAST only."};
        if (guard_cells < 4)
            throw ValueError{"active_guard_cells must be an integer >= 4."};
        if (!(truncation_sigma > 0.0) || !std::isfinite(truncation_sigma))
            throw ValueError{"active_truncation_sigma must be positive."};
        if (n_snapshots < 0 || (n_snapshots > 0 && !snapshot_energies))
            throw ValueError{"invalid snapshot request"};
        m1_ctx* c = ctx;
        double energy = c->beam_energy + n.window * c->energy_spread;
        const double emin = n.emin;
        const double tol = std::exp(-0.5 * (truncation_sigma * truncation_sigma));
        std::vector<double> targets;
        for (int i = 0; i < n_snapshots; ++i) {
            if (snapshot_taken)
                snapshot_taken[i] = 0;
            const double t = snapshot_energies[i];
            if (emin <= t && t <= energy &&
                std::find(targets.begin(), targets.end(), t) == targets.end())
                targets.push_back(t);
        }
        std::sort(targets.begin(), targets.end(), std::greater<double>());
        CellBox box = initial_box(c, guard_cells);
        const CellBox full = full_box(c);
        ensure_box_buffers(c);
        hip_check(hipMemsetAsync(c->dose, 0, size_t(c->N) * sizeof(double), c->stream),
                  "dose reset");
        // the device state (m1_set_state) is cut to the initial box
        move_state(c, full, box);
        set_box_domain(c, box);
        int32_t ne = 0;
        auto record = [&](double e) {
            if (energies && ne < energies_cap)
                energies[ne] = e;
            ++ne;
        };
        const int kind = n.scheme == M1_IF_SSPRK3 ? M1_STEP_IF_SSPRK3
                         : has_sources(n)         ? M1_STEP_STRANG
                         : n.scheme == M1_FORWARD_EULER ? M1_STEP_FORWARD_EULER
                                                        : M1_STEP_SSPRK3;
        try {
            record(energy);
            size_t ti = 0;
            int32_t steps = 0;
            while (energy > emin) {
                const double ed = m1::energy_density(energy, c->beam_energy, c->energy_spread);
                const CellBox next = desired_box(c, box, guard_cells, tol, ed);
                if (!same_box(next, box)) {
                    move_state(c, box, next);
                    box = next;
                    set_box_domain(c, box);
                }
                if (boxes && steps < energies_cap) {
                    const int32_t b6[6] = {box.i0, box.i1, box.j0, box.j1, box.k0, box.k1};
                    std::memcpy(boxes + 6 * size_t(steps), b6, sizeof b6);
                }
                double dE = n.vacuum_aware_cfl ? cfl_step_vacuum(c, n, energy, box, tol)
                                               : cfl_step(c, n, energy);
                if (ti < targets.size()) {
                    const double t = targets[ti];
                    if (energy > t && t >= energy - dE)
                        dE = energy - t;
                }
                if (energy - emin < dE)
                    dE = energy - emin;
                step(c, n, kind, energy, dE, c->box_dose, nullptr);
                hip_check(m1::box::launch_scatter_dose(c->dose, c->ny, c->nz, box, c->box_dose,
                                                       c->stream),
                          "dose scatter kernel");
                energy -= dE;
                record(energy);
                ++steps;
                if (ti < targets.size()) {
                    const double t = targets[ti];
                    if (std::fabs(energy - t) <= 1e-12) {
                        if (snapshots) {
                            int slot = 0;
                            for (int i = 0; i < n_snapshots; ++i)
                                if (snapshot_energies[i] == t) {
                                    slot = i;
                                    break;
                                }
                            hip_check(m1::box::launch_copy(c->U, box, c->V1, full, c->stream),
                                      "box copy kernel");
                            hip_check(hipMemcpyAsync(snapshots + size_t(slot) * 4 * size_t(c->N),
                                                     c->V1, 4 * size_t(c->N) * sizeof(double),
                                                     hipMemcpyDeviceToHost, c->stream),
                                      "snapshot download");
                            if (snapshot_taken)
                                snapshot_taken[slot] = 1;
                        }
                        ++ti;
                    }
                }
            }
        } catch (...) {
            // leave a full-grid state and domain behind
            set_full_domain(c);
            (void)m1::box::launch_copy(c->U, box, c->work, full, c->stream);
            std::swap(c->U, c->work);
            (void)hipStreamSynchronize(c->stream);
            throw;
        }
        // embed_state (moving_active_domain.py:235-244): the final state on the grid
        move_state(c, box, full);
        set_full_domain(c);
        sync(c);
        if (n_energies)
            *n_energies = ne;
    });
}

int m1_compute_energy_step_vacuum(m1_ctx* ctx, const m1_numerics* num, double energy,
                                  double relative_tolerance, double* dE)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        const Num n = read_numerics(num);
        if (m1::is_f32(n))
            throw NotImplemented{"the vacuum-aware CFL is computed in FP64 and FP64_FAST only."};
        if (!(relative_tolerance >= 0.0))
            throw ValueError{"relative tolerance must be non-negative"};
        ensure_box_buffers(ctx);
        const double d = cfl_step_vacuum(ctx, n, energy, full_box(ctx), relative_tolerance);
        if (dE)
            *dE = d;
    });
}

int m1_spatial_operator(m1_ctx* ctx, const m1_numerics* num, double energy,
                        const double* boundary_state, double* rhs)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        const Num n = read_numerics(num);
        if (!rhs)
            throw ValueError{"rhs must not be NULL"};
        const size_t m2 = size_t(ctx->nx) * ctx->ny;
        const bool f32 = m1::is_f32(n);
        if (boundary_state)
            hip_check(hipMemcpyAsync(ctx->bnd, boundary_state, 4 * m2 * sizeof(double),
                                     hipMemcpyHostToDevice, ctx->stream),
                      "boundary upload");
        fill_step(ctx, n, energy, 0.0, false, false, *ctx->h_step);
        if (f32) {
            narrow_state(ctx);
            upload_step32(ctx);
            if (boundary_state) {
                hip_check(m1::f32::launch_narrow_state(ctx->bnd, 1.0 / ctx->alpha, ctx->bnd32,
                                                       long(4 * m2), ctx->stream),
                          "boundary narrowing");
            } else {
                m1::BoundaryArgs32 b{dims(ctx), ctx->spatial, ctx->ids, ctx->d_step, ctx->bnd32, 0};
                hip_check(m1::f32::launch_boundary(b, ctx->stream), "boundary kernel (f32)");
            }
            m1::StageArgs32 a = base_args32(ctx, n);
            a.src = ctx->U32;
            a.U = ctx->U32;
            a.out = ctx->work32;
            a.bnd = ctx->bnd32;
            {
                Timed t(ctx->prof, ctx->stream, K_RHS);
                hip_check(m1::f32::launch_stage(m1::STAGE_RHS, false, a, ctx->stream),
                          "spatial operator kernel (f32)");
            }
            hip_check(m1::f32::launch_widen_state(ctx->work32, ctx->alpha, ctx->work, 4 * ctx->N,
                                                  ctx->stream),
                      "rhs widening");
        } else {
            upload_step(ctx);
            if (!boundary_state)
                run_boundary(ctx, n, false);
            m1::StageArgs a = base_args(ctx, n);
            a.src = ctx->U;
            a.U = ctx->U;
            a.out = ctx->work;
            a.bnd = ctx->bnd;
            Timed t(ctx->prof, ctx->stream, K_RHS);
            hip_check(stage(ctx, n, m1::STAGE_RHS, false, a), "spatial operator kernel");
        }
        hip_check(hipMemcpyAsync(rhs, ctx->work, 4 * size_t(ctx->N) * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream),
                  "rhs download");
        sync(ctx);
    });
}

int m1_compute_energy_step(m1_ctx* ctx, const m1_numerics* num, double energy, double* dE)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        const Num n = read_numerics(num);
        double d;
        if (m1::is_f32(n)) {
            narrow_state(ctx);
            d = cfl_step_f32(ctx, n, energy);
        } else {
            d = cfl_step(ctx, n, energy);
        }
        if (dE)
            *dE = d;
    });
}

int m1_step(m1_ctx* ctx, const m1_numerics* num, int32_t kind, double energy, double dE,
            double* edep)
{
    return guarded(ctx, [&] {
        require_problem(ctx);
        const Num n = read_numerics(num);
        if (m1::is_f32(n)) {
            narrow_state(ctx);
            step_f32(ctx, n, kind, energy, dE, false, edep != nullptr);
            widen_state(ctx, ctx->U);
            if (edep)
                hip_check(m1::f32::launch_widen_state(ctx->edep32, ctx->alpha, ctx->edep, ctx->N,
                                                      ctx->stream),
                          "deposition widening");
        } else {
            step(ctx, n, kind, energy, dE, nullptr, edep ? ctx->edep : nullptr);
        }
        if (edep)
            hip_check(hipMemcpyAsync(edep, ctx->edep, size_t(ctx->N) * sizeof(double),
                                     hipMemcpyDeviceToHost, ctx->stream),
                      "deposition download");
        sync(ctx);
    });
}

int m1_set_profiling(m1_ctx* ctx, int32_t enable)
{
    return guarded(ctx, [&] {
        sync(ctx);
        ctx->prof.on = enable != 0;
    });
}

int m1_set_stage_form(m1_ctx* ctx, int32_t split)
{
    return guarded(ctx, [&] {
        sync(ctx);
        ctx->split_stages = split != 0;
    });
}

int32_t m1_kernel_count(void) { return K_COUNT; }

void m1_tile_shape(int32_t* tx, int32_t* ty, int32_t* tz)
{
    if (tx)
        *tx = m1::F32_TX;
    if (ty)
        *ty = m1::F32_TY;
    if (tz)
        *tz = m1::F32_TZ;
}

const char* m1_kernel_name(int32_t kernel)
{
    return m1::kernel_name(kernel);
}

int m1_kernel_stats(m1_ctx* ctx, int32_t kernel, int64_t* launches, double* total_ms)
{
    return guarded(ctx, [&] {
        if (kernel < 0 || kernel >= K_COUNT)
            throw ValueError{"kernel index out of range"};
        sync(ctx);
        if (launches)
            *launches = ctx->prof.launches[kernel];
        if (total_ms)
            *total_ms = ctx->prof.total_ms[kernel];
    });
}

int m1_tile_stats(m1_ctx* ctx, int64_t* evaluated, int64_t* full)
{
    return guarded(ctx, [&] {
        if (evaluated)
            *evaluated = ctx->tiles_evaluated;
        if (full)
            *full = ctx->tiles_full;
    });
}

int m1_reset_stats(m1_ctx* ctx)
{
    return guarded(ctx, [&] {
        sync(ctx);
        ctx->tiles_evaluated = 0;
        ctx->tiles_full = 0;
        ctx->prof.reset();
    });
}

}  // extern "C"
