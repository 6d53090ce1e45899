// engine.hip -- batched ensembles of periodic boxes (see engine.hpp) and its C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "engine.hpp"
#include "kernels.hpp"
#include "stage_bodies.hpp"

namespace fhdx {

constexpr int EB = 256;

__device__ inline void dec3(long t, int nx, int ny, int& i, int& j, int& k)
{
    i = (int)(t % nx);
    long r = t / nx;
    j = (int)(r % ny);
    k = (int)(r / ny);
}

__device__ inline NoiseKey inst_key(const Inst& in)
{
    NoiseKey K;
    K.k0 = (uint32_t)in.seed;
    K.k1 = (uint32_t)(in.seed >> 32);
    K.step_lo = (uint32_t)in.step;
    K.field_hi = ((uint32_t)(in.step >> 32)) << 8;
    K.instance = 0;  // the same key as the single-box path: alone or batched, same noise
    return K;
}

__device__ inline double* stage_in(const Inst& in, int st) { return st == 0 ? in.U : (st == 1 ? in.U1 : in.U2); }
__device__ inline double* stage_out(const Inst& in, int st) { return st == 0 ? in.U1 : (st == 1 ? in.U2 : in.U); }

// Workgroup b runs on XCD b % 8.  XCD x takes the contiguous work-item range
// [x*nb/8, (x+1)*nb/8), so the chunks of one instance (and the z neighbours a chunk reads) stay
// on one XCD and its L2.
__device__ inline unsigned xcd_item(unsigned b, unsigned nb)
{
    const unsigned xcd = b & 7, r = b >> 3, per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + r;
}

#define EB_CELL_PROLOGUE                                       \
    const WorkItem wi = W[xcd_item(blockIdx.x, gridDim.x)];    \
    const Inst& in = I[wi.inst];                               \
    const long t = (long)wi.chunk * EB + threadIdx.x;          \
    if (t >= in.ncell || in.step >= in.step_end) return;       \
    int i, j, k;                                               \
    dec3(t, in.P.n[0], in.P.n[1], i, j, k);                    \
    const Grid G = in.G;

template <int NS>
__global__ void __launch_bounds__(EB) kb_cells(const Inst* __restrict__ I, const WorkItem* __restrict__ W, int st)
{
    EB_CELL_PROLOGUE
    cell_body<NS>(in.P, G.fld(stage_in(in, st)), G.fld(in.R), in.err, i, j, k, G.off(i, j, k));
}

__global__ void __launch_bounds__(EB) kb_nodes(const Inst* __restrict__ I, const WorkItem* __restrict__ W, double w)
{
    EB_CELL_PROLOGUE
    node_body<true>(in.P, G.fld(in.R), G.fld(in.N), G.fld(in.R), G.fld(in.R), inst_key(in), w, i, j, k,
                    G.off(i, j, k));
}

template <int NS>
__global__ void __launch_bounds__(EB) kb_faces(const Inst* __restrict__ I, const WorkItem* __restrict__ W, int st,
                                               double w)
{
    constexpr int nv = NS + 5;
    EB_CELL_PROLOGUE
    const unsigned cb = G.off(i, j, k);
    const Fld Us = G.fld(stage_in(in, st)), R = G.fld(in.R), N = G.fld(in.N);
    const NoiseKey K = inst_key(in);  // noise evaluated in place (no A/B arrays)
    double f[nv];
    face_flux<NS, 0, true>(in.P, Us, R, N, R, R, K, w, in.volinv, in.dtinv, in.dt, i, j, k, cb, f);
    store_face<NS, 0>(in.P, G.fld(in.F[0]), f, i, j, k, cb);
    face_flux<NS, 1, true>(in.P, Us, R, N, R, R, K, w, in.volinv, in.dtinv, in.dt, i, j, k, cb, f);
    store_face<NS, 1>(in.P, G.fld(in.F[1]), f, i, j, k, cb);
    face_flux<NS, 2, true>(in.P, Us, R, N, R, R, K, w, in.volinv, in.dtinv, in.dt, i, j, k, cb, f);
    store_face<NS, 2>(in.P, G.fld(in.F[2]), f, i, j, k, cb);
}

// REC: also write the next stage's cell records (stages 1 and 2), so kb_cells runs once a step.
template <int NS, bool REC>
__global__ void __launch_bounds__(EB) kb_update(const Inst* __restrict__ I, const WorkItem* __restrict__ W, int st)
{
    EB_CELL_PROLOGUE
    const Fld rn = G.fld(in.R);
    // prim is not written here: engine_read / fhdx_engine_state compute it from U on request
    update_body<NS>(in.P, G.fld(stage_out(in, st)), G.fld(in.U), G.fld(stage_in(in, st)), G.fld(in.F[0]),
                    G.fld(in.F[1]), G.fld(in.F[2]), in.dt, st, G.fld(in.prim), 0, i, j, k, G.off(i, j, k),
                    REC ? &rn : nullptr, in.err);
}

// prim (valid cells and ghost images) of one instance from its state U: conservedToPrimitive
// as the single-box step applies it after its last stage, evaluated only when prim is read.
template <int NS>
__global__ void __launch_bounds__(EB) kb_prim(DevParams P, Grid G, const double* U, double* prim, long ncell)
{
    const long t = (long)blockIdx.x * EB + threadIdx.x;
    if (t >= ncell) return;
    int i, j, k;
    dec3(t, P.n[0], P.n[1], i, j, k);
    const unsigned cb = G.off(i, j, k);
    const Fld u = G.fld(const_cast<double*>(U));
    double c[NS + 5], pr[2 * NS + 6];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) c[n] = u(cb, n);
    cons_to_prim_cell<NS>(P, c, pr);
    store_cell<2 * NS + 6, true>(P, G.fld(prim), [&](int n) { return pr[n]; }, i, j, k, cb);
}

static int compute_prim(Engine* e, const EngineSlot& s)
{
    const Inst& d = s.desc;
    const unsigned nb = (unsigned)((d.ncell + EB - 1) / EB);
    switch (s.p.nspecies) {
        case 2: kb_prim<2><<<nb, EB, 0, e->stream>>>(d.P, d.G, d.U, d.prim, d.ncell); break;
        case 3: kb_prim<3><<<nb, EB, 0, e->stream>>>(d.P, d.G, d.U, d.prim, d.ncell); break;
        case 4: kb_prim<4><<<nb, EB, 0, e->stream>>>(d.P, d.G, d.U, d.prim, d.ncell); break;
        case 5: kb_prim<5><<<nb, EB, 0, e->stream>>>(d.P, d.G, d.U, d.prim, d.ncell); break;
        case 6: kb_prim<6><<<nb, EB, 0, e->stream>>>(d.P, d.G, d.U, d.prim, d.ncell); break;
        case 7: kb_prim<7><<<nb, EB, 0, e->stream>>>(d.P, d.G, d.U, d.prim, d.ncell); break;
        case 8: kb_prim<8><<<nb, EB, 0, e->stream>>>(d.P, d.G, d.U, d.prim, d.ncell); break;
        default: return FHDX_ERR_UNSUPPORTED;
    }
    FHDX_HIP_TRY(hipGetLastError());
    FHDX_HIP_TRY(hipStreamSynchronize(e->stream));
    return FHDX_OK;
}

// Periodic ghost fill of one instance's state at admission; from then on every kernel that
// writes a cell array also writes the cell's ghost images (store_cell).
__global__ void __launch_bounds__(EB) kb_fill_state(Grid G, double* U, int nc, int g, int n0, int n1, int n2,
                                                    long nghost)
{
    const long t = (long)blockIdx.x * EB + threadIdx.x;
    if (t < nghost) ghost_body(G.view(U), nc, g, n0, n1, n2, t);
}

__global__ void kb_advance(Inst* I, const int* act, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        Inst& in = I[act[t]];
        if (in.step < in.step_end) in.step += 1;
    }
}

// ---- host ---------------------------------------------------------------------------------

static long cells_of(const fhdx_params& p) { return (long)p.n_cells[0] * p.n_cells[1] * p.n_cells[2]; }
static long ghosts_of(const fhdx_params& p)
{
    const long g = p.ngc;
    return (long)(p.n_cells[0] + 2 * g) * (p.n_cells[1] + 2 * g) * (p.n_cells[2] + 2 * g) - cells_of(p);
}
static long grown_of(const fhdx_params& p) { return cells_of(p) + ghosts_of(p); }

// Every array of an instance has the grown cell layout (one Grid per instance).
static int instance_comps(const fhdx_params& p)
{
    const int NS = p.nspecies, nv = 5 + NS, npv = 6 + 2 * NS;
    const RecLayout L{NS};
    return 3 * nv + npv + L.ncomp() + NODE_NCOMP + 3 * nv;
}

uint64_t engine_bytes_needed(const fhdx_params& p) { return 8ull * instance_comps(p) * grown_of(p); }

static int alloc_comps(EngineSlot& s, const Grid& G, double*& p, int nc)
{
    const size_t n = (size_t)G.sn * nc;
    void* ptr = nullptr;
    if (hipMalloc(&ptr, n * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc of an ensemble instance failed");
        return FHDX_ERR_NOMEM;
    }
    if (hipMemset(ptr, 0, n * sizeof(double)) != hipSuccess) return FHDX_ERR_HIP;
    s.blocks.push_back(ptr);
    p = (double*)ptr;
    return FHDX_OK;
}

static void free_slot(EngineSlot& s)
{
    for (void* ptr : s.blocks) (void)hipFree(ptr);
    s.blocks.clear();
    s.used = false;
    s.bytes = 0;
}

int engine_create(int max_inst, uint64_t budget, Engine** out)
{
    if (!out || max_inst <= 0) return FHDX_ERR_ARG;
    Engine* e = new Engine();
    e->max_inst = max_inst;
    e->budget = budget;
    e->slots.resize(max_inst);
    hipError_t r = hipMalloc((void**)&e->d_inst, sizeof(Inst) * max_inst);
    if (r == hipSuccess) r = hipMemset(e->d_inst, 0, sizeof(Inst) * max_inst);
    if (r == hipSuccess) r = hipMalloc((void**)&e->d_err, sizeof(int) * max_inst);
    if (r == hipSuccess) r = hipMemset(e->d_err, 0, sizeof(int) * max_inst);
    if (r == hipSuccess) r = hipMalloc((void**)&e->d_active, sizeof(int) * max_inst);
    if (r == hipSuccess) r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (r != hipSuccess) {
        set_error(std::string("engine setup failed: ") + hipGetErrorString(r));
        engine_destroy(e);
        return FHDX_ERR_HIP;
    }
    *out = e;
    return FHDX_OK;
}

static void drop_graph(Engine* e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

static void drop_tables(Engine* e)
{
    for (int g = 0; g < kEngineGroups; ++g) {
        if (e->d_cells[g]) (void)hipFree(e->d_cells[g]);
        e->d_cells[g] = nullptr;
        e->ncells_items[g] = 0;
    }
}

int engine_destroy(Engine* e)
{
    if (!e) return FHDX_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    drop_tables(e);
    for (EngineSlot& s : e->slots) free_slot(s);
    if (e->d_inst) (void)hipFree(e->d_inst);
    if (e->d_err) (void)hipFree(e->d_err);
    if (e->d_active) (void)hipFree(e->d_active);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return FHDX_OK;
}

int engine_add(Engine* e, const fhdx_params& p, double dt, uint64_t seed, uint64_t step0, uint64_t nsteps,
               const double* cu_init, const double* uniform, int* slot)
{
    if (!e || !slot || (!cu_init && !uniform) || !(dt > 0.)) {
        set_error("engine_add: bad argument (null slot/state or dt <= 0)");
        return FHDX_ERR_ARG;
    }
    int r = fhdx_params_check(&p);
    if (r) return r;
    if (p.nreaction > 0) {
        set_error("ensemble instances with chemistry (nreaction > 0) are not supported");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (p.transport_type == 2) {
        set_error("ensemble instances with Valk-Waldmann transport (transport_type 2) are not supported "
                  "(fhdx_rk3_step and fhdx_rk3_step_ref run it)");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (p.bc_vel_lo[0] != -1 || p.bc_vel_lo[1] != -1 || p.bc_vel_lo[2] != -1) {
        set_error("ensemble instances are periodic boxes (walls: fhdx_rk3_step_ref)");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (!fused_box_ok(p)) {
        set_error("ensemble instances need at least max(4, 2 ngc) cells per direction (thinner boxes: "
                  "fhdx_rk3_step_ref)");
        return FHDX_ERR_UNSUPPORTED;
    }
    const uint64_t need = engine_bytes_needed(p);
    if (e->used + need > e->budget) {
        set_error("ensemble memory budget exhausted");
        return FHDX_ERR_NOMEM;
    }
    int sid = -1;
    for (int q = 0; q < e->max_inst; ++q)
        if (!e->slots[q].used) {
            sid = q;
            break;
        }
    if (sid < 0) {
        set_error("no free ensemble slot");
        return FHDX_ERR_BUSY;
    }
    // the previous step of the stream may still read descriptors of this slot's old owner
    FHDX_HIP_TRY(hipStreamSynchronize(e->stream));
    EngineSlot& s = e->slots[sid];
    s = EngineSlot();
    s.used = true;
    s.p = p;
    Inst& d = s.desc;
    d.P = make_dev_params(p);
    d.G = grid_of(d.P);
    const int NS = p.nspecies, nv = 5 + NS;
    const RecLayout L{NS};
    const Grid& G = d.G;
    if (!r) r = alloc_comps(s, G, d.U, nv);
    if (!r) r = alloc_comps(s, G, d.U1, nv);
    if (!r) r = alloc_comps(s, G, d.U2, nv);
    if (!r) r = alloc_comps(s, G, d.prim, 6 + 2 * NS);
    if (!r) r = alloc_comps(s, G, d.R, L.ncomp());
    if (!r) r = alloc_comps(s, G, d.N, NODE_NCOMP);
    for (int q = 0; q < 3 && !r; ++q) r = alloc_comps(s, G, d.F[q], nv);
    if (r) {
        free_slot(s);
        return r;
    }
    d.dt = dt;
    d.volinv = 1. / (p.dx[0] * p.dx[1] * p.dx[2]);
    d.dtinv = 1. / dt;
    d.seed = seed;
    d.step = step0;
    d.step_end = nsteps ? step0 + nsteps : ~0ull;
    d.err = e->d_err + sid;
    d.ncell = cells_of(p);
    d.nghost = ghosts_of(p);
    s.bytes = need;
    e->used += need;
    if (cu_init) {
        FHDX_HIP_TRY(hipMemcpyAsync(d.U, cu_init, sizeof(double) * grown_of(p) * nv, hipMemcpyDefault, e->stream));
    } else {
        for (int c = 0; c < nv && !r; ++c) r = fill_value(d.U + c * G.sn, G.sn, uniform[c], e->stream);
        if (r) return r;
    }
    kb_fill_state<<<(unsigned)((d.nghost + EB - 1) / EB), EB, 0, e->stream>>>(G, d.U, nv, p.ngc, p.n_cells[0],
                                                                               p.n_cells[1], p.n_cells[2], d.nghost);
    FHDX_HIP_TRY(hipGetLastError());
    FHDX_HIP_TRY(hipMemcpyAsync(e->d_inst + sid, &d, sizeof(Inst), hipMemcpyHostToDevice, e->stream));
    FHDX_HIP_TRY(hipMemsetAsync(e->d_err + sid, 0, sizeof(int), e->stream));
    FHDX_HIP_TRY(hipStreamSynchronize(e->stream));
    e->dirty = true;
    *slot = sid;
    return FHDX_OK;
}

int engine_remove(Engine* e, int sid)
{
    if (!e || sid < 0 || sid >= e->max_inst || !e->slots[sid].used) return FHDX_ERR_ARG;
    FHDX_HIP_TRY(hipStreamSynchronize(e->stream));
    e->used -= e->slots[sid].bytes;
    free_slot(e->slots[sid]);
    e->dirty = true;
    return FHDX_OK;
}

static int rebuild(Engine* e)
{
    drop_graph(e);
    drop_tables(e);
    std::vector<int> act;
    for (int g = 0; g < kEngineGroups; ++g) {
        std::vector<WorkItem> cells;
        for (int q = 0; q < e->max_inst; ++q) {
            const EngineSlot& s = e->slots[q];
            if (!s.used || s.p.nspecies != g + 2) continue;
            const long nc = (s.desc.ncell + EB - 1) / EB;
            for (long c = 0; c < nc; ++c) cells.push_back({q, (int)c});
            act.push_back(q);
        }
        e->ncells_items[g] = (int)cells.size();
        if (!cells.empty()) {
            FHDX_HIP_TRY(hipMalloc((void**)&e->d_cells[g], sizeof(WorkItem) * cells.size()));
            FHDX_HIP_TRY(hipMemcpy(e->d_cells[g], cells.data(), sizeof(WorkItem) * cells.size(), hipMemcpyHostToDevice));
        }
    }
    e->nactive = (int)act.size();
    if (!act.empty())
        FHDX_HIP_TRY(hipMemcpy(e->d_active, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice));
    e->dirty = false;
    return FHDX_OK;
}

template <int NS>
static void enqueue_group(Engine* e, int g, const double ws[3])
{
    hipStream_t s = e->stream;
    const unsigned nc = (unsigned)e->ncells_items[g];
    if (!nc) return;
    const Inst* I = e->d_inst;
    const WorkItem* Wc = e->d_cells[g];
    // ghost images come with every store (store_cell): no fill launches inside the step
    kb_cells<NS><<<nc, EB, 0, s>>>(I, Wc, 0);
    for (int st = 0; st < 3; ++st) {
        kb_nodes<<<nc, EB, 0, s>>>(I, Wc, ws[st]);
        kb_faces<NS><<<nc, EB, 0, s>>>(I, Wc, st, ws[st]);
        if (st < 2) kb_update<NS, true><<<nc, EB, 0, s>>>(I, Wc, st);
        else kb_update<NS, false><<<nc, EB, 0, s>>>(I, Wc, st);
    }
}

static int enqueue_step(Engine* e)
{
    double ws[3];
    stage_weights(ws);
    enqueue_group<2>(e, 0, ws);
    enqueue_group<3>(e, 1, ws);
    enqueue_group<4>(e, 2, ws);
    enqueue_group<5>(e, 3, ws);
    enqueue_group<6>(e, 4, ws);
    enqueue_group<7>(e, 5, ws);
    enqueue_group<8>(e, 6, ws);
    if (e->nactive) kb_advance<<<(e->nactive + 255) / 256, 256, 0, e->stream>>>(e->d_inst, e->d_active, e->nactive);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int engine_step(Engine* e, int nsteps, bool use_graph)
{
    if (!e || nsteps < 0) return FHDX_ERR_ARG;
    int r;
    if (e->dirty && (r = rebuild(e))) return r;
    if (e->nactive == 0 || nsteps == 0) return FHDX_OK;
    if (use_graph) {
        if (!e->exec) {
            FHDX_HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            r = enqueue_step(e);
            hipGraph_t gph = nullptr;
            hipError_t ce = hipStreamEndCapture(e->stream, &gph);
            if (r) return r;
            if (ce != hipSuccess) {
                set_error(std::string("graph capture failed: ") + hipGetErrorString(ce));
                return FHDX_ERR_HIP;
            }
            e->graph = gph;
            FHDX_HIP_TRY(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        }
        for (int q = 0; q < nsteps; ++q) FHDX_HIP_TRY(hipGraphLaunch(e->exec, e->stream));
    } else {
        for (int q = 0; q < nsteps; ++q)
            if ((r = enqueue_step(e))) return r;
    }
    for (EngineSlot& s : e->slots)
        if (s.used) s.desc.step = std::min(s.desc.step + (uint64_t)nsteps, s.desc.step_end);
    return FHDX_OK;
}

int engine_sync(Engine* e)
{
    FHDX_HIP_TRY(hipStreamSynchronize(e->stream));
    return FHDX_OK;
}

int engine_errors(Engine* e, int* flags)
{
    FHDX_HIP_TRY(hipStreamSynchronize(e->stream));
    FHDX_HIP_TRY(hipMemcpy(flags, e->d_err, sizeof(int) * e->max_inst, hipMemcpyDeviceToHost));
    FHDX_HIP_TRY(hipMemset(e->d_err, 0, sizeof(int) * e->max_inst));
    return FHDX_OK;
}

int engine_read(Engine* e, int slot, double* cu_out, double* prim_out)
{
    if (slot < 0 || slot >= e->max_inst || !e->slots[slot].used) return FHDX_ERR_ARG;
    const EngineSlot& s = e->slots[slot];
    const size_t pts = (size_t)grown_of(s.p);
    FHDX_HIP_TRY(hipStreamSynchronize(e->stream));
    if (prim_out) {
        int r = compute_prim(e, s);
        if (r) return r;
    }
    if (cu_out)
        FHDX_HIP_TRY(hipMemcpy(cu_out, s.desc.U, sizeof(double) * pts * (5 + s.p.nspecies), hipMemcpyDefault));
    if (prim_out)
        FHDX_HIP_TRY(
            hipMemcpy(prim_out, s.desc.prim, sizeof(double) * pts * (6 + 2 * s.p.nspecies), hipMemcpyDefault));
    return FHDX_OK;
}

}  // namespace fhdx

// ---- C ABI -----------------------------------------------------------------------------

using namespace fhdx;

struct fhdx_engine {
    Engine* e;
};

extern "C" {

uint64_t fhdx_engine_bytes_needed(const fhdx_params* p) { return p ? engine_bytes_needed(*p) : 0; }

int fhdx_engine_create(int max_instances, uint64_t mem_budget_bytes, fhdx_engine** out)
{
    if (!out) return FHDX_ERR_ARG;
    Engine* e = nullptr;
    int r = engine_create(max_instances, mem_budget_bytes, &e);
    if (r) return r;
    *out = new fhdx_engine{e};
    return FHDX_OK;
}

int fhdx_engine_destroy(fhdx_engine* h)
{
    if (!h) return FHDX_OK;
    int r = engine_destroy(h->e);
    delete h;
    return r;
}

int fhdx_engine_add(fhdx_engine* h, const fhdx_params* p, double dt, uint64_t seed, uint64_t step0,
                    uint64_t nsteps, const double* cu_init, int* slot)
{
    if (!h || !p) return FHDX_ERR_ARG;
    return engine_add(h->e, *p, dt, seed, step0, nsteps, cu_init, nullptr, slot);
}

int fhdx_engine_add_uniform(fhdx_engine* h, const fhdx_params* p, double dt, uint64_t seed, uint64_t step0,
                            uint64_t nsteps, const double* values, int* slot)
{
    if (!h || !p || !values) return FHDX_ERR_ARG;
    return engine_add(h->e, *p, dt, seed, step0, nsteps, nullptr, values, slot);
}

int fhdx_engine_remove(fhdx_engine* h, int slot) { return h ? engine_remove(h->e, slot) : FHDX_ERR_ARG; }

int fhdx_engine_step(fhdx_engine* h, int nsteps, int use_graph)
{
    return h ? engine_step(h->e, nsteps, use_graph != 0) : FHDX_ERR_ARG;
}

int fhdx_engine_sync(fhdx_engine* h) { return h ? engine_sync(h->e) : FHDX_ERR_ARG; }

int fhdx_engine_state(fhdx_engine* h, int slot, fhdx_fab* cu, fhdx_fab* prim, uint64_t* step, int* done)
{
    if (!h || slot < 0 || slot >= h->e->max_inst || !h->e->slots[slot].used) return FHDX_ERR_ARG;
    const EngineSlot& s = h->e->slots[slot];
    const fhdx_params& p = s.p;
    const int g = p.ngc;
    auto fill = [&](fhdx_fab* f, double* a, int nc) {
        if (!f) return;
        f->p = a;
        for (int d = 0; d < 3; ++d) {
            f->lo[d] = -g;
            f->hi[d] = p.n_cells[d] - 1 + g;
        }
        f->ncomp = nc;
        f->pad = 0;
    };
    fill(cu, s.desc.U, 5 + p.nspecies);
    fill(prim, s.desc.prim, 6 + 2 * p.nspecies);
    if (prim) {
        FHDX_HIP_TRY(hipStreamSynchronize(h->e->stream));
        int r = compute_prim(h->e, s);
        if (r) return r;
    }
    if (step) *step = s.desc.step;
    if (done) *done = s.desc.step >= s.desc.step_end;
    return FHDX_OK;
}

int fhdx_engine_read(fhdx_engine* h, int slot, double* cu_out, double* prim_out)
{
    return h ? engine_read(h->e, slot, cu_out, prim_out) : FHDX_ERR_ARG;
}

int fhdx_engine_errors(fhdx_engine* h, int* flags)
{
    if (!h || !flags) return FHDX_ERR_ARG;
    return engine_errors(h->e, flags);
}

uint64_t fhdx_engine_bytes_used(fhdx_engine* h) { return h ? h->e->used : 0; }

}  // extern "C"
