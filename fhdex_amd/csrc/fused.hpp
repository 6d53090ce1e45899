// fused.hpp -- the fast RK3 step (fhdx_rk3_step / fhdx_rk3_step_injected).
#pragma once

#include <vector>

#include "common.hpp"

namespace fhdx {

// Per-cell record written once per stage (after conservedToPrimitive and the transport
// coefficients) and read by the node and face kernels; SoA components of a ghost-padded FAB.
struct RecLayout {
    int ns;
    static constexpr int U = 0, V = 1, W = 2, T = 3, P = 4, ETA = 5, KAP = 6, ZETA = 7, Y = 8;
    __host__ __device__ int X() const { return 8 + ns; }
    __host__ __device__ int YN() const { return 8 + 2 * ns; }    // clamped, renormalised Y
    __host__ __device__ int CHI() const { return 8 + 3 * ns; }
    __host__ __device__ int D() const { return 8 + 4 * ns; }
    __host__ __device__ int ncomp() const { return 8 + 4 * ns + ns * ns; }
};

// Node values: corner velocity gradients [derivative dir][component] (flux.cpp:1684-1703),
// visccorn (:1705) and the four stochastic edge terms sqrt(mu)*rancorn of the node.  The twelve
// edge viscosities a face sums (flux.cpp:117-135, 481-499, 847-865) use only four distinct
// node-relative summation orders: x faces all use order 0; y faces 0 (pp, pm) and 1 (mp, mm);
// z faces 0 (pp), 1 (mp), 2 (pm, the eta(k-2)*T(k) form of :857-858) and 3 (mm).  Computing
// each order once per node gives the reference's bits with 4 square roots per node instead of
// 12 per cell.
constexpr int NODE_VC = 9, NODE_S = 10, NODE_NCOMP = 14;

// Kernel groups the fast path can time with HIP events (fhdx_ctx_timing_*).
enum KernelKind { KT_NOISE = 0, KT_CELLS, KT_NODES, KT_FACES, KT_UPDATE, KT_GHOSTS, KT_COUNT };
extern const char* const kKernelKindNames[KT_COUNT];

// When enabled, brackets each kernel group with events recorded on the launch stream; the
// pairs are resolved (and the events recycled) when the totals are read.
struct Timing {
    bool on = false;
    struct Pair {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pair> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[KT_COUNT] = {};
    long launches[KT_COUNT] = {};
    int cur = -1;
    hipEvent_t cur_a = nullptr;
    hipEvent_t get();
    void begin(int kind, hipStream_t s);
    void end(hipStream_t s);
    int resolve();  // waits for every pending pair
    void reset();
    void destroy();
};

// Device workspace owned by a context for the fast path (allocated on first use).  Every
// array has the layout of a cell FAB with ngc ghost layers (component-major); face and node
// arrays hold the lower face / node of each valid cell there (face n and node n of a periodic
// box are face 0 and node 0).
struct FusedWorkspace {
    bool ready = false;
    std::vector<void*> blocks;
    double *cup = nullptr, *cup2 = nullptr, *rec = nullptr, *node = nullptr;
    double* F[3] = {nullptr, nullptr, nullptr};   // face fluxes (nvars) per direction
    double* sA[3] = {nullptr, nullptr, nullptr};  // Philox (or copied caller) noise of the step
    double* sB[3] = {nullptr, nullptr, nullptr};
    double *rA = nullptr, *rB = nullptr;
    // launch shape (fhdx_ctx_autotune picks it by timing the candidates on the device)
    int tw_max = 64;           // tile width limit in x (16, 32, 64); tile height 256 / width
    int rh = 1;                // tiles per region of the region-major block order
    int kz = 1;                // z levels each tile block marches through
    bool noise_inline = true;  // Philox noise evaluated in the faces/nodes, or once into arrays
    int face_bpc = 0;          // cap on resident k_faces blocks per CU (0: none)
    bool rec_in_update = false;  // stages 1-2: k_update writes the next stage's records
    Timing timing;
};

void fused_free(FusedWorkspace& w);

// Species counts above this evaluate the step's Philox noise into arrays (k_noise) before the
// stages, whatever noise_inline says.  At NS = 8 the face kernel with the noise evaluated in
// place (256 VGPRs + 236 AGPRs, no scratch) returned values that differ from the same kernel
// reading the noise from arrays, and from the oracle; its cause was not found.  NS 6 and 7 take
// the array kernels too, so one kernel family serves every wide mixture.
constexpr int kInlineNoiseMaxNS = 5;

// The file-scope device scratch of fhdx_rk3_step_noalloc (noalloc.hip): its size, the lookup of
// a device's copy (once per device, with that device current; fhdx_ctx_create runs it) and the
// pointer found (nullptr before the lookup).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
int noalloc_scratch_init(int device);
double* noalloc_scratch(int device);

// Device bytes of the fast path's workspace for this box with the noise evaluated in place:
// two stage states, the cell records, the node terms and three face-flux arrays (plus the
// step's noise arrays for species counts above kInlineNoiseMaxNS).
size_t fused_workspace_bytes(const DevParams& P);

// One RK3 step: cu advanced in place, prim recomputed, ghosts of both filled.  Noise from
// Philox (key K) unless sA is non-null (caller's A/B sets).
int fused_rk3_step(const fhdx_params& p, const DevParams& P, FusedWorkspace& w, fhdx_fab cu, fhdx_fab prim,
                   double dt, const NoiseKey& K, const fhdx_fab* sA, const fhdx_fab* sB, const fhdx_fab* rA,
                   const fhdx_fab* rB, int* d_err, hipStream_t s);

// The same step (Philox noise in place) with its workspace laid out in `base` (`bytes` long)
// instead of allocated, and the launch shape of `shape`.  Allocates nothing; every workspace
// value a kernel reads is written earlier in the same step, so `base` needs no clearing.
int fused_rk3_step_in(const fhdx_params& p, const DevParams& P, const FusedWorkspace& shape, double* base,
                      size_t bytes, fhdx_fab cu, fhdx_fab prim, double dt, const NoiseKey& K, int* d_err,
                      hipStream_t s);

}  // namespace fhdx
