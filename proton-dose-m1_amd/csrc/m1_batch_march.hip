// Host side of the batched-beamlet march (m1_batch_* in include/m1march.h).
//
// Each beamlet is the energy march of energy_march (numerics/energy.py:
// 1497-1644) on its own box of the phantom: its own energy E_b and CFL step,
// its own inflow, its own tile set.  One iteration of the driver advances
// every beamlet that is still marching by one of its energy steps:
//   1. one CFL launch over the tiles of all active boxes (per-box maxima,
//      read back together with the size of this step's tile list);
//   2. per-beamlet step scalars on the host (the single-box code, per box);
//   3. one boundary launch and three stage launches over the compacted
//      (box, tile) list;
//   4. per-box tile sets of the next step (mask, compaction, zeroing).
// The host synchronises once per iteration (step 1's read-back).
//
// Tile-skip network (optional, m1_batch_set_skip_model): every `horizon`
// iterations the network sees the features of every tile and marks tiles it
// expects to stay empty; for the next `horizon` steps the mask drops those
// of them that only the dilation would add.  Every `judge_every` iterations
// the step is also taken on the full box from the same state, and a beamlet
// whose step dose differs by more than judge_tol of that step's largest dose
// increment falls back to the plain mask for the rest of its march.
// Recording (m1_batch_record) collects the network's training samples from
// plain runs: the features of the tiles the dilation adds, labelled 1 when
// the tile stays at or below the threshold for the next `horizon` steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "m1_batch_steps.h"
#include "m1_host.h"
#include "m1_mlp.h"

using namespace m1;

namespace {
std::string g_batch_create_error;
}  // namespace

namespace {
// The batched kernel sets, one per compiled tile shape (the default first).
#define M1_BATCH_TABLE_ENTRY(a, b, c) &M1_BATCH_TABLE_NAME(a, b, c),
const BatchKernels& (*const kKernelSets[])() = {M1_BATCH_SHAPES(M1_BATCH_TABLE_ENTRY)};
constexpr int kKernelSetCount = int(sizeof(kKernelSets) / sizeof(kKernelSets[0]));
}  // namespace

namespace {

// Host threads for the per-beamlet step scalars, which sit between the CFL
// read-back and the next launches.  parallel_for(n, f) runs f(0..n-1) on the
// calling thread and the workers and returns when all are done; an exception
// of any f is rethrown in the caller.
class HostPool {
public:
    HostPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const int n = int(std::min(hw > 1 ? hw - 1 : 0u, 15u));
        for (int i = 0; i < n; ++i)
            workers_.emplace_back([this] { work(); });
    }
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
        }
        wake_.notify_all();
        for (auto& t : workers_)
            t.join();
    }
    void parallel_for(int n, const std::function<void(int)>& f)
    {
        if (workers_.empty() || n < 16) {
            for (int i = 0; i < n; ++i)
                f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> l(mu_);
            job_ = &f;
            n_ = n;
            next_.store(0);
            busy_ = int(workers_.size());
            error_ = nullptr;
            ++gen_;
        }
        wake_.notify_all();
        run();
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [this] { return busy_ == 0; });
        job_ = nullptr;
        if (error_)
            std::rethrow_exception(error_);
    }

private:
    void run()
    {
        for (int i; (i = next_.fetch_add(1)) < n_;) {
            try {
                (*job_)(i);
            } catch (...) {
                std::lock_guard<std::mutex> l(mu_);
                if (!error_)
                    error_ = std::current_exception();
            }
        }
    }
    void work()
    {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(mu_);
                wake_.wait(l, [&] { return stop_ || gen_ != seen; });
                if (stop_)
                    return;
                seen = gen_;
            }
            run();
            std::lock_guard<std::mutex> l(mu_);
            if (--busy_ == 0)
                done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable wake_, done_;
    const std::function<void(int)>* job_ = nullptr;
    int n_ = 0;
    std::atomic<int> next_{0};
    int busy_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
    std::exception_ptr error_;
};

}  // namespace

struct m1_batch {
    int device = 0;
    const BatchKernels* K = &kKernelSets[0]();  // kernels of the tile shape in use
    hipStream_t stream = nullptr;
    std::string err;

    bool have_problem = false;
    int px = 0, py = 0, pz = 0;
    double dx = 0, dy = 0, dz = 0;
    std::vector<MaterialData> mats;
    std::vector<m1_beamlet> bl;
    std::vector<BoxDesc> boxes;
    std::vector<double> alpha;
    std::vector<uint32_t> outside_mats;  // per box: materials of the phantom outside it
    long total_cells = 0, total_cols = 0, max_box_cells = 0;
    int total_tiles = 0, max_cols = 0;

    BoxDesc* d_box = nullptr;
    BoxStep* d_step = nullptr;
    int* d_tile_box = nullptr;
    uint8_t* d_tile_mats = nullptr;
    // W1, W2: the stencil inputs of stages 1 and 2, formed by stages 0 and 1
    float *W1 = nullptr, *W2 = nullptr;
    float *U = nullptr, *V1 = nullptr, *V2 = nullptr, *dose_hi = nullptr, *dose_lo = nullptr,
          *bnd = nullptr, *d_tile_max = nullptr, *d_tile_face = nullptr;
    uint8_t* ids = nullptr;
    double* spatial = nullptr;
    uint8_t* d_flags[2] = {nullptr, nullptr};
    int* d_list = nullptr;
    int* d_count = nullptr;  // LIST_HEADER_ALLOC ints: length, stage queues, chunk counts
    bool persistent = false; // stage kernels on a fixed grid with XCD queues
    bool cfl_all_tiles = false;  // CFL over every tile of the boxes (else the live list)
    HostPool pool;               // host threads of the per-beamlet step scalars
    int persistent_grid = 0; // resident workgroups of a stage kernel
    int cus = 0;             // compute units of the device
    bool use_graphs = false; // capture plain steps as HIP graphs
    std::string graph_error; // why the last capture failed (empty: none)
    int64_t graph_steps = 0; // steps of the last march run as a graph
    int *d_gate = nullptr, *h_gate = nullptr;  // network gate of a graph step
    StepCoef* d_coef = nullptr;
    StepCoef* d_cfl = nullptr;
    StepCoef32* d_coef32 = nullptr;
    unsigned long long* d_red = nullptr;
    unsigned long long* d_boxmats = nullptr;  // per box: materials of its vacuum tiles
    double* d_phantom = nullptr;
    double* d_box_dose = nullptr;

    StepCoef* h_coef = nullptr;
    StepCoef* h_cfl = nullptr;
    StepCoef32* h_coef32 = nullptr;
    BoxStep* h_step = nullptr;
    unsigned long long* h_red = nullptr;
    unsigned long long* h_boxmats = nullptr;
    int* h_count = nullptr;

    float* d_tile_rho = nullptr;     // mean density per tile
    float* d_tile_sums = nullptr;    // sum phi, sum |J| per tile (stage 2)
    uint8_t* d_pred = nullptr;       // network: tiles expected to stay empty
    uint8_t* d_ring = nullptr;       // tiles only the dilation adds
    uint8_t* d_cand = nullptr;       // ring tiles saved for labelling
    float* d_wmax = nullptr;         // running max of tile_max
    float* d_feat = nullptr;         // features saved for labelling
    // skip network and its judge
    const m1_mlp* model = nullptr;
    float decision = 0.5f, judge_tol = 1e-3f;
    int horizon = 8, judge_every = 64, infer_every = 4;
    int guard_cells = 4;  // TILE_GUARD (4) or 0 (m1_batch_set_skip_guard)
    bool cfl_tail_floor = false;  // m1_batch_set_cfl_tail_floor
    std::vector<uint8_t> rejected;
    std::vector<int32_t> rollback_to;  // step a rejected beamlet restarted from (-1: none)
    // two checkpoints per box (state and dose before a judged step)
    float *ck_U[2] = {nullptr, nullptr}, *ck_hi[2] = {nullptr, nullptr},
          *ck_lo[2] = {nullptr, nullptr};
    float *jf_hi = nullptr, *jf_lo = nullptr, *js_hi = nullptr, *js_lo = nullptr;
    uint8_t* d_flags_full = nullptr;
    int* d_list_full = nullptr;
    int* d_count_full = nullptr;
    unsigned *d_jdiff = nullptr, *d_jfull = nullptr, *h_jdiff = nullptr, *h_jfull = nullptr;
    uint8_t *d_sel = nullptr, *h_sel = nullptr;  // judge flags | save selectors | roll-back
    int64_t judged = 0, rejections = 0;
    // recording
    bool recording = false, record_full = false;
    int record_horizon = 8;
    int record_cap = 0;
    float *d_rx = nullptr, *d_ry = nullptr;
    int* d_rcount = nullptr;

    // quasi-steady materials (m1_batch_set_quasi_steady)
    uint32_t qs_mask = 0;
    int qs_iters = 0;
    double qs_cfl = 0.5;
    StepCoef32* d_relax32 = nullptr;  // per box: the relaxation's coefficients
    uint8_t* d_air = nullptr;         // per tile: quasi-steady tiles to relax
    int* d_air_list = nullptr;
    int* d_air_count = nullptr;
    // 4 x cells each: the sweeps' results, alternately (zero outside the
    // tiles listed for relaxation, so that a sweep reads the zero state of
    // the quasi-steady cells of tiles out of the live set from them)
    float* d_relax_tmp[2] = {nullptr, nullptr};
    double qs_tol = 0.0;              // per-box convergence of the sweeps (0: fixed count)
    uint8_t* d_relax_done = nullptr;  // 2 x boxes: converged before a sweep / after it
    unsigned* d_relax_change = nullptr;  // 3 x boxes: the sweeps' largest changes (float bits)
    long tiles_with_mat[MAXMAT] = {};  // tiles holding material m
    int progress_every = 0;  // m1_batch_set_progress: a status line every N device steps
    // device-side energy stepping (m1_batch_set_device_steps)
    bool device_steps = false;
    DevMaterial* d_mats = nullptr;  // material tables on the device
    double* d_tables = nullptr;
    BeamletConst* d_bc = nullptr;
    BeamletState* d_bst = nullptr;
    DevCounters* d_ctr = nullptr;
    StepStatus* h_ring = nullptr;   // pinned: status of the last RING steps
    StepStatus* d_status = nullptr; // its device address
    static constexpr int RING = 4;
    hipEvent_t ring_ev[RING] = {};

    Profiler prof;
    double host_ms = 0.0;
    bool marched = false;

    void free_model_buffers();
    void free_record_buffers();

    void free_problem();
    ~m1_batch();
};

void m1_batch::free_model_buffers()
{
    void* dev[] = {ck_U[0], ck_U[1], ck_hi[0], ck_hi[1], ck_lo[0], ck_lo[1], jf_hi, jf_lo,
                   js_hi, js_lo, d_flags_full, d_list_full, d_count_full, d_jdiff, d_jfull,
                   d_sel};
    for (void* p : dev)
        if (p)
            (void)hipFree(p);
    for (void* p : {(void*)h_jdiff, (void*)h_jfull, (void*)h_sel})
        if (p)
            (void)hipHostFree(p);
    for (int k = 0; k < 2; ++k)
        ck_U[k] = ck_hi[k] = ck_lo[k] = nullptr;
    jf_hi = jf_lo = js_hi = js_lo = nullptr;
    d_flags_full = d_sel = h_sel = nullptr;
    d_list_full = d_count_full = nullptr;
    d_jdiff = d_jfull = h_jdiff = h_jfull = nullptr;
}

void m1_batch::free_record_buffers()
{
    for (void* p : {(void*)d_rx, (void*)d_ry, (void*)d_rcount})
        if (p)
            (void)hipFree(p);
    d_rx = d_ry = nullptr;
    d_rcount = nullptr;
    record_cap = 0;
}

void m1_batch::free_problem()
{
    free_model_buffers();
    free_record_buffers();
    for (void* p : {(void*)d_tile_rho, (void*)d_tile_sums, (void*)d_pred, (void*)d_ring,
                    (void*)d_cand, (void*)d_wmax, (void*)d_feat})
        if (p)
            (void)hipFree(p);
    d_tile_rho = d_tile_sums = d_wmax = d_feat = nullptr;
    d_pred = d_ring = d_cand = nullptr;
    void* dev[] = {d_box,  d_step,   d_tile_box, d_tile_mats, U,         V1,       V2, W1, W2,
                   dose_hi, dose_lo, bnd,        d_tile_max,  d_tile_face, ids,       spatial,  d_flags[0],
                   d_flags[1], d_list, d_count,  d_coef,      d_cfl,     d_coef32, d_red,
                   d_boxmats, d_phantom, d_box_dose};
    for (void* p : dev)
        if (p)
            (void)hipFree(p);
    for (void* p : {(void*)d_mats, (void*)d_tables, (void*)d_bc, (void*)d_bst, (void*)d_ctr,
                    (void*)d_relax32, (void*)d_air, (void*)d_air_list, (void*)d_air_count,
                    (void*)d_relax_tmp[0], (void*)d_relax_tmp[1], (void*)d_relax_done,
                    (void*)d_relax_change})
        if (p)
            (void)hipFree(p);
    d_relax32 = nullptr;
    d_relax_tmp[0] = d_relax_tmp[1] = nullptr;
    d_relax_done = nullptr;
    d_relax_change = nullptr;
    d_air = nullptr;
    d_air_list = d_air_count = nullptr;
    d_mats = nullptr;
    d_tables = nullptr;
    d_bc = nullptr;
    d_bst = nullptr;
    d_ctr = nullptr;
    void* host[] = {h_coef, h_cfl, h_coef32, h_step, h_red, h_boxmats, h_count};
    for (void* p : host)
        if (p)
            (void)hipHostFree(p);
    d_box = nullptr;
    d_step = nullptr;
    d_tile_box = nullptr;
    d_tile_mats = nullptr;
    U = V1 = V2 = W1 = W2 = dose_hi = dose_lo = bnd = d_tile_max = d_tile_face = nullptr;
    ids = nullptr;
    spatial = nullptr;
    d_flags[0] = d_flags[1] = nullptr;
    d_list = d_count = nullptr;
    d_coef = d_cfl = nullptr;
    d_coef32 = nullptr;
    d_red = d_boxmats = nullptr;
    d_phantom = d_box_dose = nullptr;
    h_coef = h_cfl = nullptr;
    h_coef32 = nullptr;
    h_step = nullptr;
    h_red = h_boxmats = nullptr;
    h_count = nullptr;
    have_problem = false;
    marched = false;
}

m1_batch::~m1_batch()
{
    (void)hipSetDevice(device);
    if (stream)
        (void)hipStreamSynchronize(stream);
    free_problem();
    prof.release();
    if (h_ring)
        (void)hipHostFree(h_ring);
    for (hipEvent_t& e : ring_ev)
        if (e)
            (void)hipEventDestroy(e);
    if (d_gate)
        (void)hipFree(d_gate);
    if (h_gate)
        (void)hipHostFree(h_gate);
    if (stream)
        (void)hipStreamDestroy(stream);
}

namespace {

template <class F>
int guarded(m1_batch* c, F&& f)
{
    if (!c)
        return M1_ERR_STATE;
    return run_guarded(c->device, c->err, f);
}

void sync(m1_batch* c)
{
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (c->prof.on)
        c->prof.collect();
}

BeamParams beam_params(const m1_beamlet& b)
{
    return {b.beam.energy, b.beam.energy_spread, b.beam.angular_spread};
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
}

// z = 0 tiles of a box stay evaluated while the inflow spectrum, anywhere in
// the next two maximal steps, exceeds the threshold relative to its peak.
int inflow_on(const m1_beamlet& b, const Num& n, double energy_next)
{
    const double lo = energy_next - 2.0 * n.max_step;
    const double e = std::min(std::max(b.beam.energy, lo), energy_next);
    const double d = (e - b.beam.energy) / b.beam.energy_spread;
    return std::exp(-0.5 * d * d) > n.tile_threshold ? 1 : 0;
}

// Material tables of the problem on the device (one buffer of doubles, and
// the DevMaterial records pointing into it).
void upload_tables(m1_batch* c)
{
    std::vector<double> buf;
    std::vector<DevMaterial> recs(c->mats.size());
    auto put = [&](const Table& t, DevTable& d, std::vector<std::array<long, 3>>& offs) {
        std::array<long, 3> o{-1, -1, -1};
        o[0] = long(buf.size());
        buf.insert(buf.end(), t.E.begin(), t.E.end());
        o[1] = long(buf.size());
        buf.insert(buf.end(), t.v.begin(), t.v.end());
        if (!t.slope.empty()) {
            o[2] = long(buf.size());
            buf.insert(buf.end(), t.slope.begin(), t.slope.end());
        }
        d.n = int(t.E.size());
        d.pad = 0;
        offs.push_back(o);
    };
    std::vector<std::array<long, 3>> offs;
    for (size_t m = 0; m < c->mats.size(); ++m) {
        put(c->mats[m].sp, recs[m].sp, offs);
        put(c->mats[m].tr, recs[m].tr, offs);
        put(c->mats[m].nu, recs[m].nu, offs);
        recs[m].density = c->mats[m].density;
        recs[m].radiation_length = c->mats[m].radiation_length;
    }
    dev_alloc(&c->d_tables, buf.size(), "hipMalloc material tables");
    hip_check(hipMemcpy(c->d_tables, buf.data(), buf.size() * sizeof(double),
                        hipMemcpyHostToDevice),
              "table upload");
    size_t k = 0;
    for (auto& r : recs)
        for (DevTable* d : {&r.sp, &r.tr, &r.nu}) {
            const auto& o = offs[k++];
            d->E = c->d_tables + o[0];
            d->v = c->d_tables + o[1];
            d->slope = o[2] >= 0 ? c->d_tables + o[2] : nullptr;
        }
    dev_alloc(&c->d_mats, recs.size(), "hipMalloc material records");
    hip_check(hipMemcpy(c->d_mats, recs.data(), recs.size() * sizeof(DevMaterial),
                        hipMemcpyHostToDevice),
              "table upload");
}

// The host's ValueError / NonFinite for the first error of a device march.
[[noreturn]] void throw_device_error(const m1_batch* c, const DevError& e)
{
    if (e.code == DEVERR_NONFINITE)
        throw NonFinite{"non-finite M1 state detected by the CFL reduction"};
    if (e.code == DEVERR_SLOPES)
        throw ValueError{"PCHIP interpolation needs nodal slopes"};
    const char* what[3] = {"stopping-power table", "Wentzel table", "nuclear table"};
    const int tab = e.table >= 0 && e.table < 3 ? e.table : 0;
    const int m = e.material >= 0 && e.material < int(c->mats.size()) ? e.material : 0;
    const Table& t = tab == 0 ? c->mats[size_t(m)].sp
                              : (tab == 1 ? c->mats[size_t(m)].tr : c->mats[size_t(m)].nu);
    char buf[256];
    std::snprintf(buf, sizeof buf, "Energy %.17g MeV outside %s range [%.17g, %.17g] MeV.",
                  e.energy, what[tab], t.E.front(), t.E.back());
    throw ValueError{buf};
}

}  // namespace

extern "C" {

int m1_batch_create(m1_batch** out, int device)
{
    if (!out) {
        g_batch_create_error = "out must not be NULL";
        return M1_ERR_VALUE;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_batch_create_error = "no HIP device is available";
        return M1_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        g_batch_create_error = "device ordinal out of range";
        return M1_ERR_VALUE;
    }
    m1_batch* c = new m1_batch;
    c->device = device;
    try {
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        int cus = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
                  "hipDeviceGetAttribute");
        c->cus = cus;
        const int per_cu = c->K->blocks_per_cu();
        c->persistent_grid = std::max(XCD_QUEUES, (cus * per_cu) / XCD_QUEUES * XCD_QUEUES);
        dev_alloc(&c->d_gate, 1, "hipMalloc network gate");
        host_alloc(&c->h_gate, 1, "hipHostMalloc");
        *c->h_gate = 0;
        host_alloc(&c->h_ring, m1_batch::RING, "hipHostMalloc step status");
        std::memset(c->h_ring, 0, m1_batch::RING * sizeof(StepStatus));
        hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_status), c->h_ring, 0),
                  "hipHostGetDevicePointer");
        for (hipEvent_t& e : c->ring_ev)
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    } catch (const HipError& e) {
        g_batch_create_error = e.msg;
        delete c;
        return M1_ERR_HIP;
    }
    *out = c;
    return M1_OK;
}

void m1_batch_destroy(m1_batch* batch) { delete batch; }

const char* m1_batch_last_error(const m1_batch* batch)
{
    return batch ? batch->err.c_str() : g_batch_create_error.c_str();
}

int m1_batch_set_problem(m1_batch* c, int32_t nx, int32_t ny, int32_t nz, double dx, double dy,
                         double dz, const int32_t* ids, int32_t n_materials,
                         const m1_material* materials, int32_t n_beamlets,
                         const m1_beamlet* beamlets, const double* inflow_spatial)
{
    return guarded(c, [&] {
        if (nx <= 0 || ny <= 0 || nz <= 0)
            throw ValueError{"Geometry resolution must contain positive integers."};
        if (!(dx > 0 && dy > 0 && dz > 0))
            throw ValueError{"cell sizes must be positive"};
        if (n_materials < 1 || n_materials > MAXMAT)
            throw ValueError{"between 1 and 8 materials are supported"};
        if (n_beamlets < 1)
            throw ValueError{"a batch needs at least one beamlet"};
        if (!ids || !materials || !beamlets || !inflow_spatial)
            throw ValueError{"ids, materials, beamlets and inflow must not be NULL"};
        const long P = long(nx) * ny * nz;
        for (long i = 0; i < P; ++i)
            if (ids[i] < 0 || ids[i] >= n_materials)
                throw ValueError{"material id out of range"};
        std::vector<MaterialData> mats;
        for (int m = 0; m < n_materials; ++m)
            mats.push_back(read_material(materials[m], m));

        std::vector<BoxDesc> boxes(static_cast<size_t>(n_beamlets));
        std::vector<double> alpha(static_cast<size_t>(n_beamlets));
        long cells = 0, cols = 0, max_cells = 0;
        long tiles = 0;
        int max_cols = 0;
        for (int b = 0; b < n_beamlets; ++b) {
            const m1_beamlet& q = beamlets[b];
            if (q.nx <= 0 || q.ny <= 0 || q.nz <= 0)
                throw ValueError{"beamlet " + std::to_string(b) + ": empty box"};
            if (q.nz < 4)
                throw ValueError{"WENO-Z inflow extrapolation requires at least four z cells."};
            if (q.x0 < 0 || q.y0 < 0 || q.x0 + q.nx > nx || q.y0 + q.ny > ny || q.nz > nz)
                throw ValueError{"beamlet " + std::to_string(b) + ": box outside the phantom"};
            if (q.beam.energy_spread == 0.0)
                throw ValueError{"Zero energy spread is not supported by the Gaussian inflow."};
            if (!std::isfinite(q.beam.energy) || !std::isfinite(q.beam.energy_spread) ||
                !std::isfinite(q.beam.angular_spread))
                throw ValueError{"Beam parameters must be finite."};
            BoxDesc& d = boxes[size_t(b)];
            d.nx = q.nx;
            d.ny = q.ny;
            d.nz = q.nz;
            d.ntx = (q.nx + c->K->tx - 1) / c->K->tx;
            d.nty = (q.ny + c->K->ty - 1) / c->K->ty;
            d.ntz = (q.nz + c->K->tz - 1) / c->K->tz;
            d.tile_off = int(tiles);
            d.pad = 0;
            d.cell_off = cells;
            d.col_off = cols;
            const long n = long(q.nx) * q.ny * q.nz;
            const long m2 = long(q.nx) * q.ny;
            double pk = 0.0;
            for (long i = 0; i < m2; ++i) {
                const double v = inflow_spatial[cols + i];
                if (!(std::isfinite(v) && v >= 0.0))
                    throw ValueError{"inflow profile values must be finite and non-negative"};
                pk = std::max(pk, v);
            }
            alpha[size_t(b)] = march_alpha(pk, beam_params(q));
            cells += n;
            cols += m2;
            tiles += long(d.ntx) * d.nty * d.ntz;
            max_cells = std::max(max_cells, n);
            max_cols = std::max(max_cols, int(m2));
            if (tiles > (1L << 30))
                throw ValueError{"batch too large"};
        }

        // Per-box material ids and per-tile material sets.
        std::vector<uint8_t> box_ids(static_cast<size_t>(cells));
        std::vector<int> tile_box(static_cast<size_t>(tiles));
        std::vector<uint8_t> tile_mats(static_cast<size_t>(tiles), 0);
        std::vector<float> tile_rho(static_cast<size_t>(tiles), 0.f);
        std::vector<int> tile_cells(static_cast<size_t>(tiles), 0);
        // Materials of the phantom outside each box: the whole-phantom march
        // holds them at the zero state, whose CFL value bounds the step.
        std::vector<long> phantom_count(MAXMAT, 0);
        for (long i = 0; i < long(nx) * ny * nz; ++i)
            phantom_count[size_t(ids[i])] += 1;
        std::vector<uint32_t> outside_mats(size_t(n_beamlets), 0);
        for (int b = 0; b < n_beamlets; ++b) {
            const m1_beamlet& q = beamlets[b];
            const BoxDesc& d = boxes[size_t(b)];
            std::vector<long> box_count(MAXMAT, 0);
            for (int x = 0; x < q.nx; ++x)
                for (int y = 0; y < q.ny; ++y) {
                    const int32_t* src = ids + (long(q.x0 + x) * ny + (q.y0 + y)) * nz;
                    uint8_t* dst = box_ids.data() + d.cell_off + (long(x) * q.ny + y) * q.nz;
                    for (int z = 0; z < q.nz; ++z) {
                        dst[z] = uint8_t(src[z]);
                        box_count[size_t(src[z])] += 1;
                        const int t = d.tile_off +
                                      ((x / c->K->tx) * d.nty + y / c->K->ty) * d.ntz + z / c->K->tz;
                        tile_mats[size_t(t)] |= uint8_t(1u << src[z]);
                        tile_rho[size_t(t)] += float(mats[size_t(src[z])].density);
                        tile_cells[size_t(t)] += 1;
                    }
                }
            for (int t = 0; t < d.ntx * d.nty * d.ntz; ++t)
                tile_box[size_t(d.tile_off + t)] = b;
            for (int m = 0; m < MAXMAT; ++m)
                if (phantom_count[size_t(m)] > box_count[size_t(m)])
                    outside_mats[size_t(b)] |= 1u << m;
        }
        for (size_t t = 0; t < tile_rho.size(); ++t) {
            if (tile_cells[t] > 0)
                tile_rho[t] /= float(tile_cells[t]);
        }
        long with_mat[MAXMAT] = {};
        for (uint8_t tm : tile_mats)
            for (int m = 0; m < MAXMAT; ++m)
                with_mat[m] += (tm >> m) & 1;

        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        c->free_problem();
        c->px = nx;
        c->py = ny;
        c->pz = nz;
        c->dx = dx;
        c->dy = dy;
        c->dz = dz;
        c->mats = std::move(mats);
        c->bl.assign(beamlets, beamlets + n_beamlets);
        c->boxes = boxes;
        c->alpha = alpha;
        c->outside_mats = std::move(outside_mats);
        c->total_cells = cells;
        c->total_cols = cols;
        c->total_tiles = int(tiles);
        c->max_cols = max_cols;
        c->max_box_cells = max_cells;
        std::copy(with_mat, with_mat + MAXMAT, c->tiles_with_mat);
        const size_t B = size_t(n_beamlets), T = size_t(tiles), N = size_t(cells);
        dev_alloc(&c->d_box, B, "hipMalloc boxes");
        dev_alloc(&c->d_step, B, "hipMalloc box flags");
        dev_alloc(&c->d_tile_box, T, "hipMalloc tile boxes");
        dev_alloc(&c->d_tile_mats, T, "hipMalloc tile materials");
        dev_alloc(&c->U, 4 * N, "hipMalloc batch state");
        dev_alloc(&c->V1, 4 * N, "hipMalloc batch stage 1");
        dev_alloc(&c->V2, 4 * N, "hipMalloc batch stage 2");
        dev_alloc(&c->W1, 4 * N, "hipMalloc batch stage 1 input");
        dev_alloc(&c->W2, 4 * N, "hipMalloc batch stage 2 input");
        dev_alloc(&c->dose_hi, N, "hipMalloc batch dose");
        dev_alloc(&c->dose_lo, N, "hipMalloc batch dose");
        dev_alloc(&c->bnd, 12 * size_t(cols), "hipMalloc batch boundary");
        dev_alloc(&c->d_tile_max, T, "hipMalloc tile max");
        dev_alloc(&c->d_tile_face, T * FACE_SLOTS, "hipMalloc tile face slabs");
        dev_alloc(&c->ids, N, "hipMalloc batch ids");
        dev_alloc(&c->spatial, size_t(cols), "hipMalloc batch inflow");
        dev_alloc(&c->d_flags[0], T, "hipMalloc tile set");
        dev_alloc(&c->d_flags[1], T, "hipMalloc tile set");
        dev_alloc(&c->d_list, T, "hipMalloc tile list");
        dev_alloc(&c->d_count, LIST_HEADER_ALLOC, "hipMalloc tile count");
        dev_alloc(&c->d_coef, B, "hipMalloc batch coefficients");
        dev_alloc(&c->d_cfl, B, "hipMalloc batch CFL coefficients");
        dev_alloc(&c->d_coef32, B, "hipMalloc batch coefficients (f32)");
        dev_alloc(&c->d_red, B, "hipMalloc batch CFL maxima");
        dev_alloc(&c->d_boxmats, B, "hipMalloc batch CFL materials");
        dev_alloc(&c->d_box_dose, size_t(max_cells), "hipMalloc beamlet dose");
        dev_alloc(&c->d_tile_rho, T, "hipMalloc tile density");
        dev_alloc(&c->d_tile_sums, 2 * T, "hipMalloc tile sums");
        dev_alloc(&c->d_pred, T, "hipMalloc tile predictions");
        dev_alloc(&c->d_ring, T, "hipMalloc tile ring");
        dev_alloc(&c->d_cand, T, "hipMalloc tile candidates");
        dev_alloc(&c->d_wmax, T, "hipMalloc tile window max");
        dev_alloc(&c->d_feat, T * mlp::NIN, "hipMalloc tile features");
        c->rejected.assign(B, 0);
        c->rollback_to.assign(B, -1);
        host_alloc(&c->h_coef, B, "hipHostMalloc");
        host_alloc(&c->h_cfl, B, "hipHostMalloc");
        host_alloc(&c->h_coef32, B, "hipHostMalloc");
        host_alloc(&c->h_step, B, "hipHostMalloc");
        host_alloc(&c->h_red, B, "hipHostMalloc");
        host_alloc(&c->h_boxmats, B, "hipHostMalloc");
        host_alloc(&c->h_count, 1, "hipHostMalloc");
        hip_check(hipMemcpy(c->d_box, boxes.data(), B * sizeof(BoxDesc), hipMemcpyHostToDevice),
                  "box upload");
        hip_check(hipMemcpy(c->d_tile_box, tile_box.data(), T * sizeof(int),
                            hipMemcpyHostToDevice),
                  "tile upload");
        hip_check(hipMemcpy(c->d_tile_mats, tile_mats.data(), T, hipMemcpyHostToDevice),
                  "tile upload");
        hip_check(hipMemcpy(c->d_tile_rho, tile_rho.data(), T * sizeof(float),
                            hipMemcpyHostToDevice),
                  "tile upload");
        hip_check(hipMemcpy(c->ids, box_ids.data(), N, hipMemcpyHostToDevice), "ids upload");
        hip_check(hipMemcpy(c->spatial, inflow_spatial, size_t(cols) * sizeof(double),
                            hipMemcpyHostToDevice),
                  "inflow upload");
        upload_tables(c);
        dev_alloc(&c->d_bc, B, "hipMalloc beamlet constants");
        dev_alloc(&c->d_bst, B, "hipMalloc beamlet states");
        dev_alloc(&c->d_ctr, 1, "hipMalloc step counters");
        c->have_problem = true;
    });
}

int m1_batch_march(m1_batch* c, const m1_numerics* num, double* dose, int32_t* n_steps)
{
    return guarded(c, [&] {
        if (!c->have_problem)
            throw ValueError{"no problem bound: call m1_batch_set_problem first"};
        const Num n = read_numerics(num);
        if (!is_f32(n))
            throw NotImplemented{"the batched march runs the FP32-bulk kernels (M1_FP32_BULK or "
                                 "M1_FP32_BULK_F64_CLOSURE)"};
        bool ifac;
        if (n.scheme == M1_IF_SSPRK3)
            ifac = has_sources(n);
        else if (n.scheme == M1_SSPRK3 && !has_sources(n))
            ifac = false;
        else
            throw NotImplemented{"the batched march supports IF-SSPRK3, and SSPRK3 without "
                                 "scattering or nuclear removal"};
        const int B = int(c->bl.size());
        const size_t N = size_t(c->total_cells);
        const size_t T = size_t(c->total_tiles);
        hipStream_t s = c->stream;

        const size_t nb = static_cast<size_t>(B);
        std::vector<double> E(nb), dE(nb, 0.0);
        std::vector<int32_t> steps(nb, 0);
        std::vector<int> tiles_of(nb);
        int n_active = 0;
        for (int b = 0; b < B; ++b) {
            E[size_t(b)] = c->bl[size_t(b)].beam.energy +
                           n.window * c->bl[size_t(b)].beam.energy_spread;
            const BoxDesc& d = c->boxes[size_t(b)];
            tiles_of[size_t(b)] = d.ntx * d.nty * d.ntz;
            const bool on = E[size_t(b)] > n.emin;
            c->rejected[size_t(b)] = 0;
            c->rollback_to[size_t(b)] = -1;
            c->h_step[b] = {on ? 1 : 0, 1, c->model ? 1 : 0, float(E[size_t(b)]), 0};
            n_active += on ? 1 : 0;
        }
        if ((c->model || c->recording) && !(n.tile_threshold > 0.0))
            throw ValueError{"the tile-skip network and recording need tile_threshold > 0"};
        const bool use_model = c->model != nullptr;
        const bool judging = use_model && c->judge_every > 0;
        if (judging && !c->d_sel) {
            for (int k = 0; k < 2; ++k) {
                dev_alloc(&c->ck_U[k], 4 * N, "hipMalloc checkpoint state");
                dev_alloc(&c->ck_hi[k], N, "hipMalloc checkpoint dose");
                dev_alloc(&c->ck_lo[k], N, "hipMalloc checkpoint dose");
            }
            dev_alloc(&c->jf_hi, N, "hipMalloc judge dose");
            dev_alloc(&c->jf_lo, N, "hipMalloc judge dose");
            dev_alloc(&c->js_hi, N, "hipMalloc judge dose");
            dev_alloc(&c->js_lo, N, "hipMalloc judge dose");
            dev_alloc(&c->d_flags_full, T, "hipMalloc judge tile set");
            dev_alloc(&c->d_list_full, T, "hipMalloc judge tile list");
            hip_check(hipMemset(c->d_list_full, 0, T * sizeof(int)), "judge tile list");
            dev_alloc(&c->d_count_full, LIST_HEADER_ALLOC, "hipMalloc judge tile count");
            dev_alloc(&c->d_jdiff, nb, "hipMalloc judge maxima");
            dev_alloc(&c->d_jfull, nb, "hipMalloc judge maxima");
            host_alloc(&c->h_jdiff, nb, "hipHostMalloc");
            host_alloc(&c->h_jfull, nb, "hipHostMalloc");
            dev_alloc(&c->d_sel, 3 * nb, "hipMalloc judge flags");
            host_alloc(&c->h_sel, 3 * nb, "hipHostMalloc");
        }
        // Judge bookkeeping: acc[b] is the checkpoint (0 or 1) that passed the
        // last judge, -1 for the initial state; since[b] counts the steps
        // taken after it.
        std::vector<int> acc(nb, -1), since(nb, 0), pend(nb, 0);
        std::vector<double> ckE[2] = {std::vector<double>(nb), std::vector<double>(nb)};
        std::vector<int32_t> ckS[2] = {std::vector<int32_t>(nb), std::vector<int32_t>(nb)};
        const std::vector<double> E0 = E;
        uint8_t* h_judge = c->h_sel;
        uint8_t* h_save = c->h_sel + nb;
        uint8_t* h_back = c->h_sel + 2 * nb;
        auto slices = [&](int op, const uint8_t* sel) {
            BatchSliceArgs sa{c->d_box,  sel,      B,        op,       c->U,
                              c->dose_hi, c->dose_lo, {c->ck_U[0], c->ck_U[1]},
                              {c->ck_hi[0], c->ck_hi[1]}, {c->ck_lo[0], c->ck_lo[1]},
                              c->jf_hi,  c->jf_lo, c->js_hi, c->js_lo};
            hip_check(c->K->slices(sa, int(c->max_box_cells), s), "judge slice kernel");
        };
        if (c->recording)
            hip_check(hipMemsetAsync(c->d_rcount, 0, sizeof(int), s), "sample count reset");
        hip_check(hipMemsetAsync(c->d_tile_max, 0, T * sizeof(float), s), "tile max reset");
        hip_check(hipMemsetAsync(c->d_tile_sums, 0, 2 * T * sizeof(float), s), "tile sum reset");
        hip_check(hipMemsetAsync(c->d_pred, 0, T, s), "prediction reset");
        hip_check(hipMemsetAsync(c->d_wmax, 0, T * sizeof(float), s), "window max reset");
        hip_check(hipMemsetAsync(c->d_cand, 0, T, s), "candidate reset");
        for (float* p : {c->U, c->V1, c->V2, c->W1, c->W2})
            hip_check(hipMemsetAsync(p, 0, 4 * N * sizeof(float), s), "state reset");
        hip_check(hipMemsetAsync(c->dose_hi, 0, N * sizeof(float), s), "dose reset");
        hip_check(hipMemsetAsync(c->dose_lo, 0, N * sizeof(float), s), "dose reset");
        // The first step evaluates every tile.
        int cur = 0;
        hip_check(hipMemsetAsync(c->d_flags[cur], 1, T, s), "tile set");
        // Device-side stepping: energies, CFL steps, step scalars, judge
        // decisions and box flags stay on the device (not with recording).
        const bool dev = c->device_steps && !c->recording;
        const bool qs = c->qs_mask != 0 && c->qs_iters > 0;
        if (qs && !dev)
            throw NotImplemented{"quasi-steady materials need device-side stepping "
                                 "(m1_batch_set_device_steps)"};
        if (qs && c->model)
            throw NotImplemented{"quasi-steady materials and the tile-skip network are not "
                                 "combined"};
        const bool vacuum_aware0 =
            n.vacuum_aware_cfl && n.tile_threshold > 0.0 && n.wave == M1_WAVE_EXACT;
        DevStepArgs da{};
        if (dev) {
            std::vector<BeamletConst> bc(nb);
            for (size_t b = 0; b < nb; ++b) {
                const m1_beamlet& q = c->bl[b];
                bc[b] = {q.beam.energy,  q.beam.energy_spread, q.beam.angular_spread,
                         c->alpha[b],    E[b],                 c->outside_mats[b],
                         tiles_of[b]};
            }
            hip_check(hipMemcpy(c->d_bc, bc.data(), nb * sizeof(BeamletConst),
                                hipMemcpyHostToDevice),
                      "beamlet constants upload");
            hip_check(hipMemsetAsync(c->d_ctr, 0, sizeof(DevCounters), s), "counter reset");
            hip_check(hipMemsetAsync(c->d_cfl, 0, nb * sizeof(StepCoef), s), "CFL scalar reset");
            std::memset(c->h_ring, 0, m1_batch::RING * sizeof(StepStatus));
            da.mats = c->d_mats;
            da.n_mat = int(c->mats.size());
            da.n_boxes = B;
            da.scattering = n.scattering;
            da.nuclear = n.nuclear ? 1 : 0;
            da.interp = n.interp;
            da.wave = n.wave;
            da.ifac = ifac ? 1 : 0;
            da.vacuum_aware = vacuum_aware0 ? 1 : 0;
            da.cfl_box_only = n.batch_cfl_box_only ? 1 : 0;
            da.use_model = use_model ? 1 : 0;
            da.infer_every = c->infer_every;
            da.max_step = n.max_step;
            da.cfl = n.cfl;
            da.emin = n.emin;
            da.tile_threshold = n.tile_threshold;
            da.dx = c->dx;
            da.dy = c->dy;
            da.dz = c->dz;
            da.judge_tol = c->judge_tol;
            da.bc = c->d_bc;
            da.st = c->d_bst;
            da.step = c->d_step;
            da.coef = c->d_coef;
            da.cfl_coef = c->d_cfl;
            da.coef32 = c->d_coef32;
            da.red = c->d_red;
            da.box_mats = c->d_boxmats;
            da.count = c->d_count;
            da.count_full = judging ? c->d_count_full : nullptr;
            da.sel = judging ? c->d_sel : nullptr;
            da.jdiff = judging ? c->d_jdiff : nullptr;
            da.jfull = judging ? c->d_jfull : nullptr;
            da.ctr = c->d_ctr;
            da.gate = c->d_gate;
            da.ring = c->d_status;
            da.ring_len = m1_batch::RING;
            da.judge_iter = 0;
            da.qs_mask = qs ? c->qs_mask : 0u;
            if (qs) {
                // relaxation coefficients: the pseudo-time CFL number on the
                // quasi-steady materials, the box's WENO and speed floors
                if (!c->d_relax32) {
                    dev_alloc(&c->d_relax32, nb, "hipMalloc relaxation coefficients");
                    dev_alloc(&c->d_air, T, "hipMalloc quasi-steady tiles");
                    dev_alloc(&c->d_air_list, T, "hipMalloc quasi-steady tile list");
                    dev_alloc(&c->d_air_count, LIST_HEADER_ALLOC, "hipMalloc tile count");
                    for (float*& r : c->d_relax_tmp) {
                        dev_alloc(&r, 4 * N, "hipMalloc relaxation state");
                        hip_check(hipMemsetAsync(r, 0, 4 * N * sizeof(float), s),
                                  "relaxation state clear");
                    }
                    dev_alloc(&c->d_relax_done, 2 * nb, "hipMalloc relaxation flags");
                    dev_alloc(&c->d_relax_change, 3 * nb, "hipMalloc relaxation changes");
                }
                std::vector<StepCoef32> rc(nb);
                for (size_t b = 0; b < nb; ++b) {
                    StepCoef32& f = rc[b];
                    std::memset(&f, 0, sizeof f);
                    const double al = c->alpha[b];
                    f.weno_eps = float(std::max(1.0e-12 / (al * al), 1.0e-37));
                    f.eig_eps = 1.0e-10 / al;
                    for (size_t m = 0; m < c->mats.size(); ++m) {
                        f.S0[m] = f.S1[m] = f.S2[m] = 1.f;
                        f.dEm[m] = (c->qs_mask >> m & 1u) ? float(c->qs_cfl) : 0.f;
                    }
                }
                hip_check(hipMemcpy(c->d_relax32, rc.data(), nb * sizeof(StepCoef32),
                                    hipMemcpyHostToDevice),
                          "relaxation coefficient upload");
            }
            hip_check(launch_steps_init(da, s), "step init kernel");
        } else {
            hip_check(hipMemcpyAsync(c->d_step, c->h_step, size_t(B) * sizeof(BoxStep),
                                     hipMemcpyHostToDevice, s),
                      "box flag upload");
        }
        {
            // Boxes with nothing to march leave the list at once.
            BatchMaskArgs ma{c->d_box,      c->d_step, c->d_tile_box, int(T), c->d_tile_max,
                             c->d_tile_face, c->d_flags[cur], c->d_flags[1 - cur], 0.f, 0, 0, 0, 1,
                             nullptr,       nullptr,   nullptr,   nullptr,   1};
            hip_check(c->K->mask(ma, s), "tile mask kernel");
            cur = 1 - cur;
            hip_check(c->K->compact(c->d_flags[cur], int(T), c->d_list, c->d_count, s),
                      "tile compaction kernel");
        }
        const size_t nm = c->mats.size();
        // the vacuum-aware CFL needs the tile threshold (its activity level)
        // and the exact speeds (the reference applies no masks to UNIT speeds)
        const bool vacuum_aware = vacuum_aware0;
        const bool skipping = n.tile_threshold > 0.0 && !(c->recording && c->record_full);
        int64_t iter = 0;

        const float thr = float(n.tile_threshold);
        // Graph steps: the step's launches captured once per tile-set parity
        // and replayed; they need the persistent stage launch (fixed grid,
        // list length read on the device).  Judged and recording steps run
        // as separate launches, and so does every step while kernels are
        // timed (the per-launch event pairs stay outside graphs).
        bool graphs = c->use_graphs && !c->recording && !c->prof.on &&
                      !(c->qs_mask != 0 && c->qs_iters > 0);
        // (host-stepped graphs replay the persistent launch; device-stepped
        // graphs the estimated one, its grids set before each replay)
        const bool persistent = c->persistent || (graphs && !dev);

        // Stage arguments common to every launch of this march.
        BatchStageArgs a{};
        a.box = c->d_box;
        a.tile_box = c->d_tile_box;
        a.dx = c->dx;
        a.dy = c->dy;
        a.dz = c->dz;
        a.U = c->U;
        a.V1 = c->V1;
        a.ids = c->ids;
        a.bnd = c->bnd;
        a.coef = c->d_coef32;
        a.tile_max = skipping || c->recording ? c->d_tile_max : nullptr;
        a.tile_face = skipping || c->recording ? c->d_tile_face : nullptr;
        a.tile_sums = use_model || c->recording ? c->d_tile_sums : nullptr;
        a.wave_mode = n.wave;
        a.f64_closure = n.precision == M1_FP32_BULK_F64_CLOSURE ? 1 : 0;
        // The three stages over a list; count is its host-known length (the
        // grid of the one-workgroup-per-tile launch, ignored when persistent).
        // With device-side stepping the list length is not known on the host:
        // count is then an estimate (a multiple of 8) and the launch strided.
        const bool strided = dev && !persistent;
        auto overflow_wgs = [](int count) { return std::max(64, count / 256 * 8); };
        // called after the launches of stages 0 and 1 (the quasi-steady
        // relaxation of the next stage's stencil input)
        std::function<void(int)> between_stages;
        auto stages = [&](const int* list, int* header, int count, float* dhi, float* dlo) {
            a.list = list;
            a.count = header;
            a.queue = nullptr;
            // estimated launch: `count` one-tile workgroups and a few more for
            // entries past the estimate
            a.n_list = strided ? count + overflow_wgs(count) : count;
            a.strided = strided ? count : 0;
            auto persist = [&](int stage) {
                if (persistent) {
                    a.n_list = c->persistent_grid;
                    a.queue = header + 1 + XCD_QUEUES * stage;
                }
            };
            {
                persist(0);
                a.src = c->U;
                a.out = c->V1;
                a.win = nullptr;
                a.wout = c->W1;
                a.dose_hi = a.dose_lo = nullptr;
                a.bnd_stage = 0;
                Timed t(c->prof, s, K_STAGE0);
                hip_check(c->K->stage(STAGE0, false, a, s), "stage 0 kernel (batch)");
            }
            if (between_stages)
                between_stages(0);
            {
                persist(1);
                a.src = c->V1;
                a.out = c->V2;
                a.win = c->W1;
                a.wout = c->W2;
                a.bnd_stage = 1;
                Timed t(c->prof, s, K_STAGE1);
                hip_check(c->K->stage(STAGE1, ifac, a, s), "stage 1 kernel (batch)");
            }
            if (between_stages)
                between_stages(1);
            {
                persist(2);
                a.src = c->V2;
                a.out = nullptr;
                a.win = c->W2;
                a.wout = nullptr;
                a.dose_hi = dhi;
                a.dose_lo = dlo;
                a.bnd_stage = 2;
                Timed t(c->prof, s, K_STAGE2);
                hip_check(c->K->stage(STAGE2, ifac, a, s), "stage 2 kernel (batch)");
            }
        };
        // Step coefficients and boundary traces of the step about to run.
        auto upload_step = [&]() {
            hip_check(hipMemcpyAsync(c->d_coef, c->h_coef, size_t(B) * sizeof(StepCoef),
                                     hipMemcpyHostToDevice, s),
                      "coefficient upload");
            hip_check(hipMemcpyAsync(c->d_coef32, c->h_coef32, size_t(B) * sizeof(StepCoef32),
                                     hipMemcpyHostToDevice, s),
                      "coefficient upload (f32)");
            BatchBoundaryArgs ba{c->d_box, c->d_step, B,       c->spatial,
                                 c->ids,   c->d_coef, c->bnd, ifac ? 1 : 0};
            Timed t(c->prof, s, K_BOUNDARY);
            hip_check(c->K->boundary(ba, c->max_cols, s), "boundary kernel (batch)");
        };
        // Next step's tile set from the state after this step (computed set
        // flags[k], next set flags[1 - k]); the network runs when gate (device)
        // or infer (host) says so.
        auto tail = [&](int k, const int* gate, bool infer) {
            if (!dev)
                hip_check(hipMemcpyAsync(c->d_step, c->h_step, size_t(B) * sizeof(BoxStep),
                                         hipMemcpyHostToDevice, s),
                          "box flag upload");
            if (use_model && (gate || infer)) {
                // a tile predicted to stay empty is left out for `horizon` steps
                const mlp::DeviceParams prm = mlp::device_params_of(c->model);
                BatchFeatureArgs fa{c->d_box,      c->d_step,     c->d_tile_box,   int(T),
                                    c->d_flags[k], c->d_tile_max, c->d_tile_sums, c->d_tile_face,
                                    c->d_tile_rho, nullptr,       prm.w,          prm.p,
                                    c->decision,   thr,           c->d_pred,      c->horizon,
                                    gate};
                Timed t(c->prof, s, K_MODEL);
                hip_check(c->K->features(fa, s), "tile-skip network kernel");
            }
            {
                BatchMaskArgs ma{c->d_box,      c->d_step,       c->d_tile_box,
                                 int(T),        c->d_tile_max,   c->d_tile_face, c->d_flags[k],
                                 c->d_flags[1 - k], thr,
                                 tile_radius(c->K->tx), tile_radius(c->K->ty), tile_radius(c->K->tz),
                                 skipping ? 0 : 1,
                                 use_model ? c->d_pred : nullptr,
                                 c->recording ? c->d_ring : nullptr,
                                 c->recording ? c->d_wmax : nullptr,
                                 nullptr,
                                 c->guard_cells > 0 ? 1 : 0};
                Timed t(c->prof, s, K_TILE_MASK);
                hip_check(c->K->mask(ma, s), "tile mask kernel (batch)");
                hip_check(c->K->compact(c->d_flags[1 - k], int(T), c->d_list, c->d_count,
                                                s),
                          "tile compaction kernel");
            }
            BatchZeroArgs za{c->d_box,         c->d_tile_box, int(T), c->d_flags[k],
                             c->d_flags[1 - k], c->U,         c->V1,  c->V2, c->W1, c->W2};
            if (qs) {
                za.R1 = c->d_relax_tmp[0];
                za.R2 = c->d_relax_tmp[1];
            }
            Timed t(c->prof, s, K_TILE_ZERO);
            hip_check(c->K->zero(za, s), "tile zero kernel (batch)");
        };
        // CFL maxima of every active box for the state in U over the tile set
        // flags[k], read back with the length of the list built from it.
        auto fill_cfl_host = [&]() {
            const auto t0 = std::chrono::steady_clock::now();
            c->pool.parallel_for(int(B), [&](int b) {
                if (c->h_step[b].active) {
                    fill_cfl(c->mats, beam_params(c->bl[size_t(b)]), E[size_t(b)], c->h_cfl[b]);
                    set_norm(c->alpha[size_t(b)], c->h_cfl[b]);
                }
            });
            c->host_ms += ms_since(t0);
        };
        auto cfl_launch = [&](int k) {
            if (!dev)
                hip_check(hipMemcpyAsync(c->d_cfl, c->h_cfl, size_t(B) * sizeof(StepCoef),
                                         hipMemcpyHostToDevice, s),
                          "CFL coefficient upload");
            hip_check(c->K->clear(c->d_red, B, s), "CFL reset");
            hip_check(c->K->clear(c->d_boxmats, B, s), "CFL reset");
            {
                // the live list is the compaction of flags[k] (tail)
                const bool by_list = !c->cfl_all_tiles;
                BatchCflArgs ca{c->d_box,  c->d_step,      c->d_tile_box, int(T), c->dx,
                                c->dy,     c->dz,          c->U,          c->ids, c->spatial,
                                c->d_flags[k], c->d_tile_mats, c->d_cfl, c->d_red, n.wave,
                                vacuum_aware ? 1 : 0, thr,
                                by_list ? c->d_list : nullptr, c->d_count, c->d_boxmats,
                                int(B), c->max_cols, c->cfl_tail_floor ? 1 : 0};
                Timed t(c->prof, s, K_CFL);
                hip_check(c->K->cfl(ca, s), "CFL kernel (batch)");
            }
            if (dev)
                return;  // the step kernels read the maxima on the device
            hip_check(hipMemcpyAsync(c->h_red, c->d_red, size_t(B) * sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost, s),
                      "CFL read-back");
            hip_check(hipMemcpyAsync(c->h_boxmats, c->d_boxmats,
                                     size_t(B) * sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost, s),
                      "CFL read-back");
            hip_check(hipMemcpyAsync(c->h_count, c->d_count, sizeof(int), hipMemcpyDeviceToHost,
                                     s),
                      "tile count read-back");
        };
        // Everything a step launches after its host work (graph body).
        auto whole_step = [&](int k, int n_list) {
            hip_check(hipMemcpyAsync(c->d_gate, c->h_gate, sizeof(int), hipMemcpyHostToDevice, s),
                      "network gate upload");
            upload_step();
            stages(c->d_list, c->d_count, n_list, c->dose_hi, c->dose_lo);
            tail(k, c->d_gate, false);
            cfl_launch(1 - k);
        };
        // Device-side stepping: the step's scalars (CFL step first) and the
        // boundary traces; after the stages, the energies, judge decisions,
        // box flags and next CFL scalars.
        auto setup_dev = [&](int judge_iter) {
            DevStepArgs x = da;
            x.judge_iter = judge_iter;
            {
                Timed t(c->prof, s, K_STEP_SCALARS);
                hip_check(launch_steps_scalars(x, s), "step scalar kernel");
            }
            BatchBoundaryArgs ba{c->d_box, c->d_step, B,       c->spatial,
                                 c->ids,   c->d_coef, c->bnd, ifac ? 1 : 0};
            Timed t(c->prof, s, K_BOUNDARY);
            hip_check(c->K->boundary(ba, c->max_cols, s), "boundary kernel (batch)");
        };
        auto advance_dev = [&](int judge_iter) {
            DevStepArgs x = da;
            x.judge_iter = judge_iter;
            Timed t(c->prof, s, K_STEP_ADVANCE);
            hip_check(launch_steps_advance(x, s), "step advance kernel");
        };
        // Quasi-steady materials: after each stage, qs_iters Jacobi sweeps of
        // a pseudo-time step towards div F = 0 on their cells of the next
        // stage's stencil input (W1, W2, then U; the other cells fixed), over
        // the live tiles of active boxes holding them.  The quasi-steady cells
        // thus follow the other cells at every stage (no lag of a step).
        long air_bound = 0;
        for (int m = 0; m < MAXMAT; ++m)
            if (c->qs_mask >> m & 1u)
                air_bound += c->tiles_with_mat[m];
        const int air_est = int(std::max(8L, std::min(long(T), (air_bound + 7) / 8 * 8)));
        // Sweep j reads the quasi-steady cells from the output of sweep j - 1
        // (the first from the state) and writes them to d_relax_tmp[j % 2];
        // the last sweep's output is committed to the state.  With a
        // tolerance, a box stops sweeping once a sweep changes none of its
        // quasi-steady cells by more than it (the float state is scaled to the
        // beamlet's peak inflow fluence, so the bound is relative to each
        // beamlet): its tiles then carry their cells through the remaining
        // sweeps.  The flags alternate between two halves, the changes rotate
        // through three (read, filled, cleared).
        const bool qs_conv = c->qs_tol > 0.0;
        auto relax_air = [&](int k, float* state, int bnd_stage, bool build_list) {
            const int last = (c->qs_iters - 1) % 2;
            BatchAirArgs aa{c->d_box, c->d_step, c->d_tile_box, int(T), c->d_flags[k],
                            c->d_tile_mats, c->qs_mask, c->d_air, c->d_air_list, c->d_air_count,
                            c->ids, c->d_relax_tmp[last], state};
            aa.n_boxes = int(B);
            aa.done = c->d_relax_done;
            aa.change = c->d_relax_change;
            if (build_list) {
                hip_check(c->K->air_mask(aa, s), "quasi-steady tile mask");
                hip_check(c->K->compact(c->d_air, int(T), c->d_air_list, c->d_air_count, s),
                          "quasi-steady tile list");
            }
            BatchStageArgs ra = a;
            ra.list = c->d_air_list;
            ra.count = c->d_air_count;
            ra.queue = nullptr;
            ra.strided = air_est;  // at least the tiles holding the materials: no overflow
            ra.n_list = air_est;
            ra.src = state;
            ra.U = state;
            ra.win = nullptr;
            ra.wout = nullptr;
            ra.coef = c->d_relax32;
            ra.dose_hi = ra.dose_lo = nullptr;
            ra.tile_max = nullptr;
            ra.tile_sums = nullptr;
            ra.tile_face = nullptr;
            ra.judge = nullptr;
            ra.bnd_stage = bnd_stage;
            ra.relax_tol = float(c->qs_tol);
            Timed t(c->prof, s, K_RELAX);
            if (qs_conv)
                hip_check(c->K->air_reset(aa, s), "relaxation flag reset");
            for (int it = 0; it < c->qs_iters; ++it) {
                ra.V1 = it == 0 ? nullptr : c->d_relax_tmp[(it - 1) % 2];
                ra.out = c->d_relax_tmp[it % 2];
                if (qs_conv) {
                    ra.relax_done_in = c->d_relax_done + (it % 2) * B;
                    ra.relax_done_out = c->d_relax_done + ((it + 1) % 2) * B;
                    ra.relax_change_prev = c->d_relax_change + ((it + 2) % 3) * B;
                    ra.relax_change_cur = c->d_relax_change + (it % 3) * B;
                    ra.relax_change_clear = c->d_relax_change + ((it + 1) % 3) * B;
                }
                hip_check(c->K->stage(STAGE_RELAX, false, ra, s), "relaxation kernel");
            }
            hip_check(c->K->air_commit(aa, std::min(air_est, 16384), s), "relaxation commit");
        };
        int graph_est = 8;  // the estimate a device-stepped graph is captured with
        int graph_set[2] = {0, 0};  // the estimate each graph's stage grids hold
        auto whole_step_dev = [&](int k) {
            setup_dev(0);
            stages(c->d_list, c->d_count, graph_est, c->dose_hi, c->dose_lo);
            advance_dev(0);
            tail(k, c->d_gate, false);
            cfl_launch(1 - k);
        };
        hipGraphExec_t graph[2] = {nullptr, nullptr};
        // Device-stepped graphs: the captured graphs and their stage-kernel
        // nodes, whose grids and estimates are set before each replay.
        struct StageNode {
            hipGraphNode_t node;
            int form;
            hipKernelNodeParams p;
            BatchStageArgs args;
        };
        hipGraph_t graph_src[2] = {nullptr, nullptr};
        std::vector<StageNode> stage_nodes[2];
        auto find_stage_nodes = [&](int k, hipGraph_t g) -> bool {
            size_t n = 0;
            if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess)
                return false;
            std::vector<hipGraphNode_t> nodes(n);
            if (n && hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess)
                return false;
            stage_nodes[k].clear();
            for (hipGraphNode_t nd : nodes) {
                hipGraphNodeType t;
                if (hipGraphNodeGetType(nd, &t) != hipSuccess || t != hipGraphNodeTypeKernel)
                    continue;
                hipKernelNodeParams p{};
                if (hipGraphKernelNodeGetParams(nd, &p) != hipSuccess)
                    return false;
                const int form = c->K->stage_form(p.func);
                if (form < 0)
                    continue;
                if (form == 2 || !p.kernelParams || !p.kernelParams[0])
                    return false;
                StageNode sn{nd, form, p, {}};
                std::memcpy(&sn.args, p.kernelParams[0], sizeof(BatchStageArgs));
                stage_nodes[k].push_back(sn);
            }
            // every stage as a one-tile launch and its overflow launch
            int forms[2] = {0, 0};
            for (const StageNode& sn : stage_nodes[k])
                forms[sn.form] += 1;
            return forms[0] == 3 && forms[1] == 3;
        };
        auto set_graph_grids = [&](int k, int count) -> bool {
            for (StageNode& sn : stage_nodes[k]) {
                BatchStageArgs x = sn.args;
                const int over = overflow_wgs(count);
                x.strided = count;
                x.n_list = count + over;
                hipKernelNodeParams p = sn.p;
                p.gridDim = dim3(unsigned(sn.form == 1 ? over : count));
                void* args[1] = {&x};
                p.kernelParams = args;
                p.extra = nullptr;
                if (hipGraphExecKernelNodeSetParams(graph[k], sn.node, &p) != hipSuccess)
                    return false;
            }
            return true;
        };
        auto capture = [&](int k, int n_list) -> hipGraphExec_t {
            const hipError_t b = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (b != hipSuccess) {
                c->graph_error = hipGetErrorString(b);
                (void)hipGetLastError();
                return nullptr;
            }
            std::string why;
            try {
                if (dev)
                    whole_step_dev(k);
                else
                    whole_step(k, n_list);
            } catch (const HipError& e) {
                why = e.msg;
            }
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            hipGraphExec_t ex = nullptr;
            if (why.empty() && e == hipSuccess && g)
                e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (why.empty() && e == hipSuccess && ex && dev && !find_stage_nodes(k, g))
                why = "stage-kernel nodes of the captured step not found";
            if (g && dev && why.empty() && e == hipSuccess && ex)
                graph_src[k] = g;  // kept: its nodes address the replay's updates
            else if (g)
                (void)hipGraphDestroy(g);
            if (!why.empty() || e != hipSuccess || !ex) {
                if (ex)
                    (void)hipGraphExecDestroy(ex);
                ex = nullptr;
                c->graph_error = why.empty() ? std::string(hipGetErrorString(e)) : why;
                (void)hipGetLastError();
                return nullptr;
            }
            return ex;
        };
        struct GraphGuard {
            hipGraphExec_t* g;
            hipGraph_t* src;
            ~GraphGuard()
            {
                for (int k = 0; k < 2; ++k) {
                    if (g[k])
                        (void)hipGraphExecDestroy(g[k]);
                    if (src[k])
                        (void)hipGraphDestroy(src[k]);
                }
            }
        } graph_guard{graph, graph_src};
        c->graph_steps = 0;

        // Energies after a step and next step's box flags; the judge's
        // accept / roll-back when the step was judged.  Returns roll-backs.
        auto advance = [&](int n_judge) {
            n_active = 0;
            int n_back = 0;
            for (int b = 0; b < B; ++b) {
                const size_t k = size_t(b);
                c->h_step[b].full_set = 0;
                if (n_judge)
                    h_back[b] = 0;
                if (!c->h_step[b].active)
                    continue;
                E[k] -= dE[k];
                steps[k] += 1;
                since[k] += 1;
                if (n_judge && h_judge[b]) {
                    // accepted: the pending checkpoint becomes the last good
                    // one; rejected: back to the last good checkpoint, on
                    // every tile for one step, without the network after
                    float df, fl;
                    std::memcpy(&df, &c->h_jdiff[b], sizeof df);
                    std::memcpy(&fl, &c->h_jfull[b], sizeof fl);
                    c->judged += 1;
                    if (!(df <= c->judge_tol * fl) && !(fl == 0.f && df == 0.f)) {
                        c->rejected[k] = 1;
                        c->rejections += 1;
                        h_back[b] = uint8_t(acc[k] < 0 ? 3 : 1 + acc[k]);
                        E[k] = acc[k] < 0 ? E0[k] : ckE[acc[k]][k];
                        steps[k] = acc[k] < 0 ? 0 : ckS[acc[k]][k];
                        c->rollback_to[k] = steps[k];
                        c->h_step[b].full_set = 1;
                        n_back += 1;
                    } else {
                        acc[k] = pend[k];
                        since[k] = 1;
                    }
                }
                const bool on = E[k] > n.emin;
                c->h_step[b].active = on ? 1 : 0;
                c->h_step[b].inflow_on = on ? inflow_on(c->bl[k], n, E[k]) : 0;
                c->h_step[b].model_on = use_model && !c->rejected[k] ? 1 : 0;
                c->h_step[b].energy = float(E[k]);
                n_active += on ? 1 : 0;
            }
            return n_back;
        };

        if (dev) {
            // Steps are enqueued without waiting for them: the host reads the
            // status of step k - RING (pinned memory, written by that step's
            // scalar kernel) before enqueueing step k, to stop once no beamlet
            // marches and to size the stage grids from that step's list
            // length.  Steps enqueued after the last one find every beamlet
            // done and change nothing.  Judged steps (every judge_every-th)
            // are launched one by one, the others replayed as graphs when
            // graphs are on.
            const int R = m1_batch::RING;
            auto round8 = [](long v) { return int(std::max(8L, (v + 7) / 8 * 8)); };
            const int all_tiles = round8(long(T));
            int est = all_tiles;  // the first step evaluates every tile
            int prev_count = int(T);
            if (qs)
                between_stages = [&](int done) {
                    relax_air(cur, done == 0 ? c->W1 : c->W2, done + 1, done == 0);
                };
            cfl_launch(cur);
            for (int64_t k = 0;; ++k) {
                if (k >= R) {
                    hip_check(hipEventSynchronize(c->ring_ev[k % R]), "hipEventSynchronize");
                    const volatile StepStatus* vs = c->h_ring + k % R;
                    const int st_active = vs->n_active, st_count = vs->count, st_err = vs->err;
                    if (st_err != 0) {
                        sync(c);
                        DevCounters ctr;
                        hip_check(hipMemcpy(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost),
                                  "counter read-back");
                        throw_device_error(c, ctr.err);
                    }
                    if (st_active == 0)
                        break;
                    if (c->progress_every > 0 && (k - R) % c->progress_every == 0)
                        std::fprintf(stderr, "m1 batch: energy step %lld, %d beamlets marching\n",
                                     static_cast<long long>(k - R), st_active);
                    const long mx = std::max(st_count, prev_count);
                    prev_count = st_count;
                    est = std::min(all_tiles, round8(mx + mx / 4 + 64));
                    if (c->prof.on && k % 32 == 0)
                        sync(c);
                }
                const bool jstep = judging && k % c->judge_every == c->judge_every - 1;
                bool replayed = false;
                if (graphs && !jstep) {
                    // the estimate rounded up to 1/16 - 1/8 of itself, so that most
                    // replays keep the grids of the last one
                    int q = 64;
                    while (q * 16 < est)
                        q *= 2;
                    const int est_q = std::min(all_tiles, (est + q - 1) / q * q);
                    graph_est = est_q;
                    bool set_ok = true;
                    if (!graph[cur]) {
                        graph[cur] = capture(cur, 0);
                        graph_set[cur] = est_q;
                    } else if (graph_set[cur] != est_q) {
                        set_ok = set_graph_grids(cur, est_q);
                        graph_set[cur] = est_q;
                    }
                    if (graph[cur] && !set_ok) {
                        c->graph_error = "setting the stage grids of a graph failed";
                        (void)hipGetLastError();
                        graphs = false;
                    } else if (graph[cur]) {
                        hip_check(hipGraphLaunch(graph[cur], s), "hipGraphLaunch");
                        c->graph_steps += 1;
                        replayed = true;
                    } else {
                        graphs = false;
                    }
                }
                if (!replayed) {
                    setup_dev(jstep ? 1 : 0);
                    if (!jstep) {
                        stages(c->d_list, c->d_count, est, c->dose_hi, c->dose_lo);
                    } else {
                        // as the host-stepped judge below, with the flags set by
                        // the scalar kernel and the decision in the advance kernel
                        const uint8_t* d_judge = c->d_sel;
                        const uint8_t* d_save = c->d_sel + nb;
                        slices(SLICE_SAVE, d_save);
                        slices(SLICE_RESET, d_judge);
                        BatchMaskArgs fa{c->d_box,   c->d_step, c->d_tile_box, int(T), c->d_tile_max,
                                         c->d_tile_face, c->d_flags[cur], c->d_flags_full, 0.f, 0,
                                         0, 0, 1, nullptr, nullptr, nullptr, d_judge, 1};
                        hip_check(c->K->mask(fa, s), "judge tile set");
                        hip_check(c->K->compact(c->d_flags_full, int(T), c->d_list_full,
                                                c->d_count_full, s),
                                  "judge tile list");
                        stages(c->d_list_full, c->d_count_full, all_tiles, c->jf_hi, c->jf_lo);
                        slices(SLICE_RESTORE, d_save);
                        BatchZeroArgs za{c->d_box, c->d_tile_box, int(T), c->d_flags_full,
                                         c->d_flags[cur], c->U, c->V1, c->V2, c->W1, c->W2};
                        hip_check(c->K->zero(za, s), "judge tile reset");
                        a.judge = d_judge;
                        a.judge_hi = c->js_hi;
                        a.judge_lo = c->js_lo;
                        stages(c->d_list, c->d_count, est, c->dose_hi, c->dose_lo);
                        a.judge = nullptr;
                        slices(SLICE_ADD, d_judge);
                        hip_check(hipMemsetAsync(c->d_jdiff, 0, nb * sizeof(unsigned), s),
                                  "judge reset");
                        hip_check(hipMemsetAsync(c->d_jfull, 0, nb * sizeof(unsigned), s),
                                  "judge reset");
                        BatchCompareArgs ca{c->d_box, d_judge,  B,         c->jf_hi,  c->jf_lo,
                                            c->js_hi, c->js_lo, c->d_jdiff, c->d_jfull};
                        hip_check(c->K->compare(ca, int(c->max_box_cells), s), "judge compare");
                    }
                    if (qs)
                        relax_air(cur, c->U, 2, false);
                    advance_dev(jstep ? 1 : 0);
                    if (jstep)
                        slices(SLICE_RESTORE, c->d_sel + 2 * nb);
                    tail(cur, nullptr, use_model && (k + 1) % c->infer_every == 0);
                    cfl_launch(1 - cur);
                }
                hip_check(hipEventRecord(c->ring_ev[k % R], s), "hipEventRecord");
                cur = 1 - cur;
            }
            sync(c);
            DevCounters ctr;
            hip_check(hipMemcpy(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost),
                      "counter read-back");
            if (ctr.err.code != 0)
                throw_device_error(c, ctr.err);
            std::vector<BeamletState> bst(nb);
            hip_check(hipMemcpy(bst.data(), c->d_bst, nb * sizeof(BeamletState),
                                hipMemcpyDeviceToHost),
                      "beamlet state read-back");
            for (size_t b = 0; b < nb; ++b) {
                steps[b] = bst[b].steps;
                c->rejected[b] = uint8_t(bst[b].rejected);
                c->rollback_to[b] = bst[b].rollback_to;
            }
            c->judged += int64_t(ctr.judged);
            c->rejections += int64_t(ctr.rejections);
            c->prof.tiles_evaluated += int64_t(ctr.tiles_evaluated);
            c->prof.tiles_full += int64_t(ctr.tiles_full);
        } else {
        // CFL of the first step.
        fill_cfl_host();
        cfl_launch(cur);
        sync(c);

        while (n_active > 0) {
            const int n_list = *c->h_count;
            c->prof.tiles_evaluated += n_list;

            // 1. Step scalars of every active beamlet.
            auto t0 = std::chrono::steady_clock::now();
            c->pool.parallel_for(int(B), [&](int b) {
                if (!c->h_step[b].active)
                    return;
                double mx;
                std::memcpy(&mx, &c->h_red[b], sizeof mx);
                if (!vacuum_aware) {
                    // cells at the zero state -- the box's tiles outside the
                    // live set (materials from the CFL kernel) and, in the
                    // whole-phantom march, the phantom outside the box:
                    // isotropic speeds (wave_speeds.py:37-50)
                    // (the float isotropic speed of the kernels' closure)
                    const double iso = n.wave == M1_WAVE_UNIT ? 1.0 : double(0.57735026918962576f);
                    const uint64_t zero_mats =
                        c->h_boxmats[b] |
                        (n.batch_cfl_box_only ? 0u : c->outside_mats[size_t(b)]);
                    for (int m = 0; m < int(nm); ++m)
                        if (zero_mats >> m & 1u) {
                            const double s0 = c->h_cfl[b].S0[m];
                            const double v = iso / (s0 * c->dx) + iso / (s0 * c->dy) +
                                             iso / (s0 * c->dz);
                            if (v > mx)
                                mx = v;
                        }
                }
                double d = cfl_energy_step(n, mx);
                if (E[size_t(b)] - n.emin < d)
                    d = E[size_t(b)] - n.emin;
                dE[size_t(b)] = d;
                fill_step(c->mats, beam_params(c->bl[size_t(b)]), n, E[size_t(b)], d, ifac,
                          false, c->h_coef[b]);
                set_norm(c->alpha[size_t(b)], c->h_coef[b]);
                fill_step32(nm, c->alpha[size_t(b)], c->h_coef[b], c->h_coef32[b]);
            });
            for (int b = 0; b < B; ++b)
                if (c->h_step[b].active)
                    c->prof.tiles_full += tiles_of[size_t(b)];
            c->host_ms += ms_since(t0);
            // Beamlets judged this step: every judge_every steps after their
            // checkpoint, and on their last step.
            int n_judge = 0, n_full = 0;
            if (judging)
                for (int b = 0; b < B; ++b) {
                    const size_t k = size_t(b);
                    h_judge[b] = h_save[b] = 0;
                    if (!c->h_step[b].active || !c->h_step[b].model_on)
                        continue;
                    const bool last = !(E[k] - dE[k] > n.emin);
                    if (since[k] + 1 < c->judge_every && !last)
                        continue;
                    pend[k] = acc[k] == 0 ? 1 : 0;
                    ckE[pend[k]][k] = E[k];
                    ckS[pend[k]][k] = steps[k];
                    h_judge[b] = 1;
                    h_save[b] = uint8_t(1 + pend[k]);
                    n_judge += 1;
                    n_full += tiles_of[k];
                }

            if (graphs && n_judge == 0) {
                // 2a. The host work of the whole step first, then its launches
                // as one graph (no judge result to wait for).
                (void)advance(0);
                if (n_active == 0) {
                    upload_step();
                    stages(c->d_list, c->d_count, n_list, c->dose_hi, c->dose_lo);
                    break;
                }
                iter += 1;
                *c->h_gate = use_model && iter % c->infer_every == 0 ? 1 : 0;
                fill_cfl_host();
                if (!graph[cur] && graphs)
                    graph[cur] = capture(cur, n_list);
                if (graph[cur]) {
                    hip_check(hipGraphLaunch(graph[cur], s), "hipGraphLaunch");
                    c->graph_steps += 1;
                } else {
                    graphs = false;
                    whole_step(cur, n_list);
                }
                sync(c);
                cur = 1 - cur;
                continue;
            }

            // 2b. Step by separate launches.
            upload_step();
            if (n_judge == 0) {
                stages(c->d_list, c->d_count, n_list, c->dose_hi, c->dose_lo);
            } else {
                // The judged beamlets take the step twice from the same state:
                // on every tile of their box (dose into the full pair), then
                // on the skip set with the other beamlets (dose into the skip
                // pair, then added to the dose).  State and dose before the
                // step are saved as the beamlet's pending checkpoint.
                const uint8_t* d_judge = c->d_sel;
                const uint8_t* d_save = c->d_sel + nb;
                hip_check(hipMemcpyAsync(c->d_sel, c->h_sel, 2 * nb, hipMemcpyHostToDevice, s),
                          "judge flag upload");
                slices(SLICE_SAVE, d_save);
                slices(SLICE_RESET, d_judge);
                BatchMaskArgs fa{c->d_box,   c->d_step, c->d_tile_box, int(T), c->d_tile_max,
                                 c->d_tile_face, c->d_flags[cur], c->d_flags_full, 0.f, 0, 0, 0, 1,
                                 nullptr,      nullptr,   nullptr,       d_judge, 1};
                hip_check(c->K->mask(fa, s), "judge tile set");
                hip_check(c->K->compact(c->d_flags_full, int(T), c->d_list_full,
                                                c->d_count_full, s),
                          "judge tile list");
                c->prof.tiles_evaluated += n_full;
                stages(c->d_list_full, c->d_count_full, n_full, c->jf_hi, c->jf_lo);
                slices(SLICE_RESTORE, d_save);
                BatchZeroArgs za{c->d_box, c->d_tile_box, int(T), c->d_flags_full,
                                 c->d_flags[cur], c->U, c->V1, c->V2, c->W1, c->W2};
                hip_check(c->K->zero(za, s), "judge tile reset");
                a.judge = d_judge;
                a.judge_hi = c->js_hi;
                a.judge_lo = c->js_lo;
                stages(c->d_list, c->d_count, n_list, c->dose_hi, c->dose_lo);
                a.judge = nullptr;
                slices(SLICE_ADD, d_judge);
                hip_check(hipMemsetAsync(c->d_jdiff, 0, nb * sizeof(unsigned), s),
                          "judge reset");
                hip_check(hipMemsetAsync(c->d_jfull, 0, nb * sizeof(unsigned), s),
                          "judge reset");
                BatchCompareArgs ca{c->d_box, d_judge,  B,         c->jf_hi,  c->jf_lo,
                                    c->js_hi, c->js_lo, c->d_jdiff, c->d_jfull};
                hip_check(c->K->compare(ca, int(c->max_box_cells), s), "judge compare");
                hip_check(hipMemcpyAsync(c->h_jdiff, c->d_jdiff, nb * sizeof(unsigned),
                                         hipMemcpyDeviceToHost, s),
                          "judge read-back");
                hip_check(hipMemcpyAsync(c->h_jfull, c->d_jfull, nb * sizeof(unsigned),
                                         hipMemcpyDeviceToHost, s),
                          "judge read-back");
                sync(c);
            }

            // 3. Energies and the next step's tile sets.
            const int n_back = advance(n_judge);
            if (n_back) {
                hip_check(hipMemcpyAsync(c->d_sel + 2 * nb, h_back, nb, hipMemcpyHostToDevice, s),
                          "roll-back flag upload");
                slices(SLICE_RESTORE, c->d_sel + 2 * nb);
            }
            if (n_active == 0)
                break;
            iter += 1;
            tail(cur, nullptr, use_model && iter % c->infer_every == 0);
            if (c->recording && iter % c->record_horizon == 0) {
                // label the ring saved `horizon` steps ago, then save this step's
                if (iter >= c->record_horizon) {
                    BatchLabelArgs la{int(T),     c->d_cand, c->d_feat, c->d_wmax, thr,
                                      c->d_rx,    c->d_ry,   c->d_rcount, c->record_cap};
                    hip_check(c->K->labels(la, s), "label kernel");
                }
                BatchFeatureArgs fa{c->d_box,    c->d_step,    c->d_tile_box,   int(T),
                                    c->d_flags[cur], c->d_tile_max, c->d_tile_sums, c->d_tile_face,
                                    c->d_tile_rho,
                                    c->d_feat,   nullptr,      nullptr,         0.f,
                                    thr,         nullptr,      0,               nullptr};
                hip_check(c->K->features(fa, s), "feature kernel");
                hip_check(hipMemcpyAsync(c->d_cand, c->d_ring, T, hipMemcpyDeviceToDevice, s),
                          "candidate save");
                hip_check(hipMemsetAsync(c->d_wmax, 0, T * sizeof(float), s), "window reset");
            }
            // 4. CFL of the next step.
            fill_cfl_host();
            cfl_launch(1 - cur);
            sync(c);
            cur = 1 - cur;
        }
        }  // host-side stepping

        if (dose) {
            const size_t P = size_t(c->px) * c->py * c->pz;
            if (!c->d_phantom)
                dev_alloc(&c->d_phantom, P, "hipMalloc phantom dose");
            hip_check(hipMemsetAsync(c->d_phantom, 0, P * sizeof(double), s), "dose reset");
            {
                Timed t(c->prof, s, K_DOSE);
                for (int b = 0; b < B; ++b) {
                    const m1_beamlet& q = c->bl[size_t(b)];
                    const BoxDesc& d = c->boxes[size_t(b)];
                    hip_check(c->K->scatter_dose(c->dose_hi + d.cell_off,
                                                         c->dose_lo + d.cell_off, q.nx, q.ny, q.nz,
                                                         q.x0, q.y0, c->py, c->pz, c->d_phantom, s),
                              "dose scatter kernel");
                }
            }
            hip_check(hipMemcpyAsync(dose, c->d_phantom, P * sizeof(double),
                                     hipMemcpyDeviceToHost, s),
                      "dose download");
        }
        sync(c);
        if (n_steps)
            std::memcpy(n_steps, steps.data(), size_t(B) * sizeof(int32_t));
        c->marched = true;
    });
}

int m1_batch_beamlet_dose(m1_batch* c, int32_t beamlet, double* dose)
{
    return guarded(c, [&] {
        if (!c->have_problem)
            throw ValueError{"no problem bound: call m1_batch_set_problem first"};
        if (beamlet < 0 || beamlet >= int(c->bl.size()))
            throw ValueError{"beamlet index out of range"};
        if (!dose)
            throw ValueError{"dose must not be NULL"};
        if (!c->marched)
            throw ValueError{"no march has run on this batch"};
        const BoxDesc& d = c->boxes[size_t(beamlet)];
        const long n = long(d.nx) * d.ny * d.nz;
        hip_check(c->K->box_dose(c->dose_hi + d.cell_off, c->dose_lo + d.cell_off, n,
                                         c->d_box_dose, c->stream),
                  "beamlet dose kernel");
        hip_check(hipMemcpyAsync(dose, c->d_box_dose, size_t(n) * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream),
                  "dose download");
        sync(c);
    });
}

int m1_batch_set_profiling(m1_batch* c, int32_t enable)
{
    return guarded(c, [&] {
        sync(c);
        c->prof.on = enable != 0;
    });
}

int m1_batch_kernel_stats(m1_batch* c, int32_t kernel, int64_t* launches, double* total_ms)
{
    return guarded(c, [&] {
        if (kernel < 0 || kernel >= K_COUNT)
            throw ValueError{"kernel index out of range"};
        sync(c);
        if (launches)
            *launches = c->prof.launches[kernel];
        if (total_ms)
            *total_ms = c->prof.total_ms[kernel];
    });
}

int m1_batch_tile_stats(m1_batch* c, int64_t* evaluated, int64_t* full)
{
    return guarded(c, [&] {
        if (evaluated)
            *evaluated = c->prof.tiles_evaluated;
        if (full)
            *full = c->prof.tiles_full;
    });
}

int m1_batch_reset_stats(m1_batch* c)
{
    return guarded(c, [&] {
        sync(c);
        c->prof.reset();
        c->host_ms = 0.0;
    });
}

int m1_batch_host_ms(m1_batch* c, double* setup_ms)
{
    return guarded(c, [&] {
        if (setup_ms)
            *setup_ms = c->host_ms;
    });
}

int32_t m1_batch_tile_shape_count(void) { return kKernelSetCount; }

int m1_batch_tile_shape(int32_t index, int32_t* tx, int32_t* ty, int32_t* tz)
{
    if (index < 0 || index >= kKernelSetCount)
        return M1_ERR_VALUE;
    const BatchKernels& k = kKernelSets[index]();
    if (tx)
        *tx = k.tx;
    if (ty)
        *ty = k.ty;
    if (tz)
        *tz = k.tz;
    return M1_OK;
}

int m1_batch_set_tile_shape(m1_batch* c, int32_t index)
{
    return guarded(c, [&] {
        if (index < 0 || index >= kKernelSetCount)
            throw ValueError{"tile shape index out of range"};
        if (c->have_problem)
            throw ValueError{"set the tile shape before m1_batch_set_problem"};
        c->K = &kKernelSets[index]();
        c->persistent_grid =
            std::max(XCD_QUEUES, (c->cus * c->K->blocks_per_cu()) / XCD_QUEUES * XCD_QUEUES);
    });
}

int m1_batch_set_graphs(m1_batch* c, int32_t enable)
{
    return guarded(c, [&] {
        sync(c);
        c->use_graphs = enable != 0;
        c->graph_error.clear();
    });
}

int m1_batch_graph_info(m1_batch* c, int64_t* graph_steps, const char** capture_error)
{
    return guarded(c, [&] {
        if (graph_steps)
            *graph_steps = c->graph_steps;
        if (capture_error)
            *capture_error = c->graph_error.c_str();
    });
}

int m1_batch_set_device_steps(m1_batch* c, int32_t enable)
{
    return guarded(c, [&] { c->device_steps = enable != 0; });
}

int m1_batch_set_progress(m1_batch* c, int32_t every)
{
    return guarded(c, [&] {
        if (every < 0)
            throw ValueError{"every must be >= 0"};
        c->progress_every = every;
    });
}

int m1_batch_set_quasi_steady(m1_batch* c, uint32_t material_mask, int32_t iterations,
                              double pseudo_cfl)
{
    return guarded(c, [&] {
        if (material_mask >> MAXMAT)
            throw ValueError{"material_mask names materials beyond the 8 supported"};
        if (iterations < 0 || iterations > 4096)
            throw ValueError{"iterations must lie in [0, 4096]"};
        if (!(pseudo_cfl > 0.0 && pseudo_cfl <= 1.0))
            throw ValueError{"pseudo_cfl must lie in (0, 1]"};
        c->qs_mask = material_mask;
        c->qs_iters = iterations;
        c->qs_cfl = pseudo_cfl;
    });
}

int m1_batch_set_quasi_steady_tolerance(m1_batch* c, double tolerance)
{
    return guarded(c, [&] {
        if (!(tolerance >= 0.0 && tolerance < 1.0))
            throw ValueError{"tolerance must lie in [0, 1)"};
        c->qs_tol = tolerance;
    });
}

int m1_batch_physics_probe(m1_batch* c, int32_t material, int32_t kind, int32_t scattering,
                           int32_t interpolation, const double* energies, double dE, int32_t n,
                           double* out)
{
    return guarded(c, [&] {
        if (!c->have_problem)
            throw ValueError{"no problem bound: call m1_batch_set_problem first"};
        if (material < 0 || material >= int(c->mats.size()))
            throw ValueError{"material index out of range"};
        if (kind < 0 || kind > 4)
            throw ValueError{"kind must lie in [0, 4]"};
        if (n < 0 || (n > 0 && (!energies || !out)))
            throw ValueError{"energies and out must hold n values"};
        if (n == 0)
            return;
        double *d_e = nullptr, *d_o = nullptr;
        DevError* d_err = nullptr;
        struct Free {
            void* p[3];
            ~Free()
            {
                for (void* q : p)
                    if (q)
                        (void)hipFree(q);
            }
        } guard{{nullptr, nullptr, nullptr}};
        dev_alloc(&d_e, size_t(n), "hipMalloc probe energies");
        guard.p[0] = d_e;
        dev_alloc(&d_o, size_t(n), "hipMalloc probe output");
        guard.p[1] = d_o;
        dev_alloc(&d_err, 1, "hipMalloc probe error");
        guard.p[2] = d_err;
        hip_check(hipMemcpy(d_e, energies, size_t(n) * sizeof(double), hipMemcpyHostToDevice),
                  "probe upload");
        hip_check(hipMemsetAsync(d_err, 0, sizeof(DevError), c->stream), "probe error reset");
        hip_check(launch_physics_probe(c->d_mats, material, kind, scattering, interpolation, d_e,
                                       dE, n, d_o, d_err, c->stream),
                  "physics probe kernel");
        sync(c);
        DevError e;
        hip_check(hipMemcpy(&e, d_err, sizeof e, hipMemcpyDeviceToHost), "probe read-back");
        if (e.code != 0)
            throw_device_error(c, e);
        hip_check(hipMemcpy(out, d_o, size_t(n) * sizeof(double), hipMemcpyDeviceToHost),
                  "probe read-back");
    });
}

int m1_batch_set_cfl_all_tiles(m1_batch* c, int32_t enable)
{
    return guarded(c, [&] {
        sync(c);
        c->cfl_all_tiles = enable != 0;
    });
}

int m1_batch_set_persistent(m1_batch* c, int32_t enable)
{
    return guarded(c, [&] {
        sync(c);
        c->persistent = enable != 0;
    });
}

int m1_batch_set_skip_model(m1_batch* c, const m1_mlp* model, float decision, int32_t horizon,
                            int32_t infer_every, int32_t judge_every, float judge_tol)
{
    return guarded(c, [&] {
        if (model) {
            if (!(decision > 0.f && decision < 1.f))
                throw ValueError{"decision must lie in (0, 1)"};
            if (horizon < 1 || horizon > 255)
                throw ValueError{"horizon must lie in [1, 255]"};
            if (infer_every < 1 || infer_every > horizon)
                throw ValueError{"infer_every must lie in [1, horizon]"};
            if (judge_every < 0)
                throw ValueError{"judge_every must be >= 0 (0: no judge)"};
            if (!(judge_tol > 0.f))
                throw ValueError{"judge_tol must be positive"};
            if (mlp::device_of(model) != c->device)
                throw ValueError{"the model lives on another device"};
            hip_check(mlp::finish(model), "model synchronisation");
        }
        sync(c);
        c->model = model;
        c->decision = decision;
        c->horizon = horizon;
        c->infer_every = infer_every;
        c->judge_every = judge_every;
        c->judge_tol = judge_tol;
        c->judged = c->rejections = 0;
    });
}

int m1_batch_set_cfl_tail_floor(m1_batch* c, int32_t enable)
{
    return guarded(c, [&] {
        sync(c);
        c->cfl_tail_floor = enable != 0;
    });
}

int m1_batch_set_skip_guard(m1_batch* c, int32_t guard_cells)
{
    return guarded(c, [&] {
        if (guard_cells != 0 && guard_cells != TILE_GUARD)
            throw ValueError{"guard_cells must be 0 or " + std::to_string(TILE_GUARD)};
        sync(c);
        c->guard_cells = guard_cells;
    });
}

int m1_batch_record(m1_batch* c, int32_t enable, int32_t full_domain, int32_t horizon,
                    int64_t capacity)
{
    return guarded(c, [&] {
        sync(c);
        c->free_record_buffers();
        c->recording = false;
        if (!enable)
            return;
        if (horizon < 1)
            throw ValueError{"horizon must be at least 1"};
        if (capacity < 1 || capacity > (int64_t(1) << 28))
            throw ValueError{"capacity must lie in [1, 2^28]"};
        dev_alloc(&c->d_rx, size_t(capacity) * mlp::NIN, "hipMalloc samples");
        dev_alloc(&c->d_ry, size_t(capacity), "hipMalloc labels");
        dev_alloc(&c->d_rcount, 1, "hipMalloc sample count");
        hip_check(hipMemset(c->d_rcount, 0, sizeof(int)), "sample count reset");
        c->record_cap = int(capacity);
        c->record_horizon = horizon;
        c->record_full = full_domain != 0;
        c->recording = true;
    });
}

int m1_batch_samples(m1_batch* c, float* x, float* y, int64_t capacity, int64_t* count)
{
    return guarded(c, [&] {
        if (!count)
            throw ValueError{"count must not be NULL"};
        int n = 0;
        if (c->d_rcount) {
            hip_check(hipMemcpyAsync(&n, c->d_rcount, sizeof(int), hipMemcpyDeviceToHost,
                                     c->stream),
                      "sample count read-back");
            sync(c);
        }
        n = std::min(n, c->record_cap);
        *count = n;
        const int64_t k = std::min<int64_t>(n, capacity);
        if (k > 0) {
            if (!x || !y)
                throw ValueError{"x and y must not be NULL"};
            hip_check(hipMemcpyAsync(x, c->d_rx, size_t(k) * mlp::NIN * sizeof(float),
                                     hipMemcpyDeviceToHost, c->stream),
                      "sample download");
            hip_check(hipMemcpyAsync(y, c->d_ry, size_t(k) * sizeof(float),
                                     hipMemcpyDeviceToHost, c->stream),
                      "label download");
            sync(c);
        }
    });
}

int m1_batch_skip_stats(m1_batch* c, int64_t* judged, int64_t* rejections, int32_t* rollback_to)
{
    return guarded(c, [&] {
        if (judged)
            *judged = c->judged;
        if (rejections)
            *rejections = c->rejections;
        if (rollback_to && !c->rollback_to.empty())
            std::memcpy(rollback_to, c->rollback_to.data(),
                        c->rollback_to.size() * sizeof(int32_t));
    });
}

}  // extern "C"
