// Device-side energy stepping of the batched march: per-beamlet energies,
// CFL steps, step scalars, judge decisions and box flags kept on the device,
// so that consecutive energy steps of the batch are enqueued (or replayed as
// one HIP graph per step) without a host round trip in between.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "m1_device.h"
#include "m1_physics_dev.h"

namespace m1 {

// Per beamlet, fixed for a march.
struct BeamletConst {
    double energy, energy_spread, angular_spread;
    double alpha;        // normalisation of the float state (peak inflow fluence)
    double E0;           // first energy: energy + window * spread
    uint32_t outside_mats;  // materials of the phantom outside the box
    int tiles;           // tiles of the box
};

// Per beamlet, changed by the step kernels.
struct BeamletState {
    double E;            // energy at the start of the step being set up / taken
    double dE;           // energy step of the step being taken
    double ckE[2];       // energies of the two judge checkpoints
    int steps;           // energy steps taken
    int acc;             // checkpoint that passed the last judge (-1: the start)
    int pend;            // checkpoint of a judged step in flight
    int ckS[2];          // steps of the two checkpoints
    int rejected;        // the judge rejected the network for this beamlet
    int rollback_to;     // step it restarted from (-1: none)
    int pad;
};

// Counters of the device march (one set per batch).
struct DevCounters {
    int n_active;        // beamlets still marching after the last advance
    int iter;            // energy steps enqueued so far (the step being set up)
    unsigned long long tiles_evaluated;
    unsigned long long tiles_full;
    unsigned long long judged;
    unsigned long long rejections;
    DevError err;
};

// Status of one step, written to pinned host memory by the step's scalar
// kernel: the host reads it a few steps later (no stall of the device).
struct StepStatus {
    int iter;            // step index
    int n_active;        // beamlets marching in this step
    int count;           // tiles in this step's list
    int err;             // first device error code so far
};

struct DevStepArgs {
    const DevMaterial* mats;
    int n_mat;
    int n_boxes;
    int scattering, nuclear, interp, wave, ifac;
    int vacuum_aware;        // the CFL kernels applied the vacuum-aware masks
    int cfl_box_only;        // no outside-the-box materials in the CFL bound
    int use_model;
    int infer_every;
    uint32_t qs_mask;        // quasi-steady materials: held through a step, no CFL bound
    double max_step, cfl, emin, tile_threshold;
    double dx, dy, dz;
    float judge_tol;
    const BeamletConst* bc;
    BeamletState* st;
    BoxStep* step;
    StepCoef* coef;          // scalars of the step being taken
    StepCoef* cfl_coef;      // CFL scalars at the energy the next step starts from
    StepCoef32* coef32;
    const unsigned long long* red;      // CFL maxima (ordered bits)
    const unsigned long long* box_mats; // materials of each box's vacuum tiles
    const int* count;        // list header of the step
    const int* count_full;   // judge list header
    uint8_t* sel;            // judge | save | roll-back flags, 3 x n_boxes
    const unsigned* jdiff;
    const unsigned* jfull;
    DevCounters* ctr;
    int* gate;               // network gate of the next step's tail
    StepStatus* ring;        // pinned host memory, ring_len entries
    int ring_len;
    int judge_iter;          // this step is judged (host schedule)
};

// Launchers (m1_batch_steps.hip).
hipError_t launch_steps_init(const DevStepArgs& a, hipStream_t s);
hipError_t launch_steps_scalars(const DevStepArgs& a, hipStream_t s);
hipError_t launch_steps_advance(const DevStepArgs& a, hipStream_t s);
// Device evaluation of the scalar physics for tests: out[i] for energies[i]
// of material m; kind 0 S(E), 1 transport coefficient, 2 removal
// coefficient, 3 scattering depth over [E - dE, E], 4 nuclear depth.
hipError_t launch_physics_probe(const DevMaterial* mats, int m, int kind, int model, int interp,
                                const double* energies, double dE, int n, double* out,
                                DevError* err, hipStream_t s);

}  // namespace m1
