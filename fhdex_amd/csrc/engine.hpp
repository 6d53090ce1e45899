// engine.hpp -- batched ensembles: many independent periodic boxes advanced together.
//
// Each admitted instance owns its state and workspace in HBM and a descriptor in a device
// table.  One launch per kernel kind covers every instance of a species count: each block reads
// a work item (instance, chunk of 256 cells) and runs the same per-cell/face/node bodies as the
// single-box fast path, so an instance's result is bit-identical whether it runs alone or in a
// batch.  A whole step of the current membership is captured once in a hipGraph and replayed.
#pragma once

#include <vector>

#include "fused.hpp"
#include "stage_bodies.hpp"

namespace fhdx {

// Device descriptor of one instance.  Every array has the instance's grown cell layout G; the
// noise is evaluated in place from (seed, step), so no noise arrays exist.
struct Inst {
    DevParams P;
    Grid G;
    double *U, *U1, *U2, *prim, *R, *N;
    double* F[3];
    double dt, volinv, dtinv;
    uint64_t seed;
    uint64_t step;      // next step index (noise key); advanced on the device
    uint64_t step_end;  // the instance idles once step reaches step_end
    int* err;           // per-instance error flags
    long ncell, nghost;
};

struct WorkItem {
    int inst;
    int chunk;
};

struct EngineSlot {
    bool used = false;
    fhdx_params p{};
    Inst desc{};
    std::vector<void*> blocks;
    uint64_t bytes = 0;
};

// species counts 2..FHDX_MAX_SPECIES, one work table each
constexpr int kEngineGroups = FHDX_MAX_SPECIES - 1;

struct Engine {
    int max_inst = 0;
    uint64_t budget = 0, used = 0;
    std::vector<EngineSlot> slots;
    Inst* d_inst = nullptr;
    int* d_err = nullptr;
    hipStream_t stream = nullptr;
    // work tables per species count (index ns-2): cell chunks
    WorkItem* d_cells[kEngineGroups] = {};
    int ncells_items[kEngineGroups] = {};
    int* d_active = nullptr;  // active slot list for the step counter kernel
    int nactive = 0;
    bool dirty = true;        // membership changed: rebuild tables, recapture the graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

uint64_t engine_bytes_needed(const fhdx_params& p);
int engine_create(int max_inst, uint64_t budget, Engine** out);
int engine_destroy(Engine* e);
// cu_init: the grown conserved state (host or device); when NULL, component c of every cell is
// set to uniform[c] instead.
int engine_add(Engine* e, const fhdx_params& p, double dt, uint64_t seed, uint64_t step0, uint64_t nsteps,
               const double* cu_init, const double* uniform, int* slot);
int engine_remove(Engine* e, int slot);
int engine_step(Engine* e, int nsteps, bool use_graph);
int engine_sync(Engine* e);
// Waits, then copies (and clears) the per-slot error flags into flags[max_inst].
int engine_errors(Engine* e, int* flags);
// Waits, then copies a slot's grown conserved / primitive state (either pointer may be NULL).
int engine_read(Engine* e, int slot, double* cu_out, double* prim_out);

}  // namespace fhdx
