// server.hip -- continuous-batching request loop over the ensemble engine.
//
// Callers submit ensemble runs (parameters, dt, seed, nsteps, start state) from any host
// thread.  One worker thread owns the engine: it admits queued requests -- in arrival order, or
// smallest work (cells x steps) first -- while the memory budget, a slot and an optional cap on
// the cells in flight allow, advances every admitted instance together in chunks of
// steps (new arrivals join at the next chunk boundary), copies each finished instance's result
// to host memory, frees its slot and records a completion with arrival, admission and
// completion times.  A request whose instance raises an error flag is evicted alone; the others
// continue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

using namespace fhdx;

namespace {

int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Pending {
    int64_t id;
    fhdx_request r;
    double uniform[FHDX_MAX_SPECIES + 5];
    int64_t t_arrival;
    uint64_t cells() const
    {
        return (uint64_t)r.params.n_cells[0] * r.params.n_cells[1] * r.params.n_cells[2];
    }
    double work() const { return (double)cells() * (double)r.nsteps; }
};

struct Running {
    Pending q;
    int slot;
    int64_t t_admit;
};

}  // namespace

struct fhdx_server {
    Engine* e = nullptr;
    int device = 0;
    int chunk = 8;
    bool use_graph = true;

    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Pending> incoming;          // guarded by mu
    std::deque<fhdx_completion> finished;  // guarded by mu
    int64_t next_id = 0;                   // guarded by mu
    int64_t open = 0;                      // submitted, not yet completed; guarded by mu
    bool stop = false;                     // guarded by mu
    int policy = 0;                        // FHDX_ADMIT_*; guarded by mu
    uint64_t max_active_cells = 0;         // 0: no cap; guarded by mu

    // worker thread only
    std::thread worker;
    std::deque<Pending> pending;
    std::vector<Running> running;
    std::vector<int> flags;
    std::atomic<uint64_t> steps_done{0}, cell_updates{0}, bytes_used{0};

    void complete(const Pending& q, int status, int64_t t_admit, uint64_t steps, const char* msg)
    {
        fhdx_completion c{};
        c.id = q.id;
        c.status = status;
        c.arrival_ns = q.t_arrival;
        c.admit_ns = t_admit;
        c.done_ns = now_ns();
        c.steps = steps;
        if (msg) std::strncpy(c.message, msg, sizeof(c.message) - 1);
        std::lock_guard<std::mutex> lk(mu);
        finished.push_back(c);
        --open;
        cv_done.notify_all();
    }

    void admit(int pol, uint64_t cap)
    {
        if (pol == FHDX_ADMIT_SMALLEST_FIRST)
            std::stable_sort(pending.begin(), pending.end(),
                             [](const Pending& a, const Pending& b) { return a.work() < b.work(); });
        uint64_t active = 0;
        for (const Running& r : running) active += r.q.cells();
        while (!pending.empty()) {
            Pending& q = pending.front();
            const fhdx_params& p = q.r.params;
            int rc = fhdx_params_check(&p);
            if (rc) {
                complete(q, rc, 0, 0, fhdx_last_error());
                pending.pop_front();
                continue;
            }
            const uint64_t need = engine_bytes_needed(p);
            if (need > e->budget) {
                complete(q, FHDX_ERR_NOMEM, 0, 0, "request exceeds the whole memory budget");
                pending.pop_front();
                continue;
            }
            bool free_slot = false;
            for (const EngineSlot& s : e->slots) free_slot = free_slot || !s.used;
            if (e->used + need > e->budget || !free_slot) break;  // wait for memory / a slot
            if (cap && active > 0 && active + q.cells() > cap) break;  // the GPU is full enough
            int slot = -1;
            rc = engine_add(e, p, q.r.dt, q.r.seed, 0, q.r.nsteps, q.r.cu_init, q.r.cu_init ? nullptr : q.uniform,
                            &slot);
            if (rc) {
                complete(q, rc, 0, 0, fhdx_last_error());
                pending.pop_front();
                continue;
            }
            running.push_back({q, slot, now_ns()});
            active += q.cells();
            pending.pop_front();
            bytes_used = e->used;
        }
    }

    int finish(Running& r, int status, const char* msg)
    {
        const uint64_t step = e->slots[r.slot].desc.step;
        // the result goes where the request asked for it; a request with neither buffer only
        // wants its completion (the state is dropped with the slot)
        if (status == FHDX_OK && (r.q.r.cu_out || r.q.r.prim_out)) {
            int rc = engine_read(e, r.slot, r.q.r.cu_out, r.q.r.prim_out);
            if (rc) {
                status = rc;
                msg = fhdx_last_error();
            }
        }
        const std::string m = msg ? msg : "";
        int rc = engine_remove(e, r.slot);
        bytes_used = e->used;
        complete(r.q, status, r.t_admit, step, msg ? m.c_str() : nullptr);
        return rc;
    }

    void run()
    {
        (void)hipSetDevice(device);
        flags.assign(e->max_inst, 0);
        for (;;) {
            int pol;
            uint64_t cap;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (pending.empty() && running.empty())
                    cv.wait(lk, [&] { return stop || !incoming.empty(); });
                if (stop) break;
                while (!incoming.empty()) {
                    pending.push_back(incoming.front());
                    incoming.pop_front();
                }
                pol = policy;
                cap = max_active_cells;
            }
            admit(pol, cap);
            if (running.empty()) continue;
            uint64_t chunk_steps = (uint64_t)chunk;
            for (const Running& r : running) {
                const Inst& d = e->slots[r.slot].desc;
                chunk_steps = std::min(chunk_steps, d.step_end - d.step);
            }
            uint64_t cells = 0;
            for (const Running& r : running) cells += (uint64_t)e->slots[r.slot].desc.ncell;
            int rc = engine_step(e, (int)chunk_steps, use_graph);
            if (!rc) rc = engine_errors(e, flags.data());
            if (rc) {
                // a failed launch leaves the device state unknown: fail everything admitted
                std::string m = fhdx_last_error();
                for (Running& r : running) (void)finish(r, rc, m.c_str());
                running.clear();
                continue;
            }
            steps_done += chunk_steps;
            cell_updates += cells * chunk_steps;
            std::vector<Running> keep;
            for (Running& r : running) {
                const Inst& d = e->slots[r.slot].desc;
                if (flags[r.slot]) (void)finish(r, FHDX_ERR_MASSFRAC, "mass fraction outside (0,1)");
                else if (d.step >= d.step_end) (void)finish(r, FHDX_OK, nullptr);
                else keep.push_back(r);
            }
            running.swap(keep);
        }
        // stopped: everything not finished completes as cancelled
        for (Running& r : running) (void)finish(r, FHDX_ERR_BUSY, "server stopped");
        running.clear();
        std::deque<Pending> rest;
        {
            std::lock_guard<std::mutex> lk(mu);
            rest.swap(incoming);
        }
        for (const Pending& q : pending) complete(q, FHDX_ERR_BUSY, 0, 0, "server stopped");
        for (const Pending& q : rest) complete(q, FHDX_ERR_BUSY, 0, 0, "server stopped");
        pending.clear();
    }
};

extern "C" {

int fhdx_server_create(int max_instances, uint64_t mem_budget_bytes, int chunk_steps, int use_graph,
                       fhdx_server** out)
{
    if (!out || max_instances <= 0 || chunk_steps <= 0) return FHDX_ERR_ARG;
    fhdx_server* s = new fhdx_server();
    int rc = engine_create(max_instances, mem_budget_bytes, &s->e);
    if (rc) {
        delete s;
        return rc;
    }
    (void)hipGetDevice(&s->device);
    s->chunk = chunk_steps;
    s->use_graph = use_graph != 0;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return FHDX_OK;
}

int fhdx_server_set_policy(fhdx_server* s, int policy, uint64_t max_active_cells)
{
    if (!s || (policy != FHDX_ADMIT_ARRIVAL && policy != FHDX_ADMIT_SMALLEST_FIRST)) return FHDX_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->policy = policy;
        s->max_active_cells = max_active_cells;
    }
    s->cv.notify_one();
    return FHDX_OK;
}

int fhdx_server_submit(fhdx_server* s, const fhdx_request* r, int64_t* id)
{
    if (!s || !r) return FHDX_ERR_ARG;
    if (r->nsteps < 1 || !(r->dt > 0.)) {
        set_error("fhdx_server_submit: nsteps must be >= 1 and dt > 0");
        return FHDX_ERR_ARG;
    }
    Pending q{};
    q.t_arrival = now_ns();
    q.r = *r;
    const fhdx_params& p = r->params;
    if (!r->cu_init) {
        // uniform start (InitConsVar prob_type 1): rho, zero momentum, rho*e, rho*Y_k, with
        // e = e0mix + T*cvmix summed as GetEnergy does (compressible_functions.H:1498-1511)
        if (p.nspecies < 1 || p.nspecies > FHDX_MAX_SPECIES) return FHDX_ERR_ARG;
        double cvmix = 0., e0mix = 0.;
        for (int k = 0; k < p.nspecies; ++k) {
            cvmix += r->Y0[k] * p.hcv[k];
            e0mix += r->Y0[k] * p.e0[k];
        }
        const double ie = e0mix + r->T0 * cvmix;
        q.uniform[0] = r->rho0;
        q.uniform[1] = q.uniform[2] = q.uniform[3] = 0.;
        q.uniform[4] = r->rho0 * ie;
        for (int k = 0; k < p.nspecies; ++k) q.uniform[5 + k] = r->rho0 * r->Y0[k];
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return FHDX_ERR_BUSY;
        q.id = s->next_id++;
        ++s->open;
        s->incoming.push_back(q);
    }
    s->cv.notify_one();
    if (id) *id = q.id;
    return FHDX_OK;
}

int fhdx_server_poll(fhdx_server* s, fhdx_completion* out, int max, int* n)
{
    if (!s || !n || (max > 0 && !out)) return FHDX_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    int k = 0;
    while (k < max && !s->finished.empty()) {
        out[k++] = s->finished.front();
        s->finished.pop_front();
    }
    *n = k;
    return FHDX_OK;
}

int fhdx_server_wait(fhdx_server* s, double timeout_s, int64_t* open)
{
    if (!s) return FHDX_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    bool idle = s->cv_done.wait_until(lk, until, [&] { return s->open == 0; });
    if (open) *open = s->open;
    return idle ? FHDX_OK : FHDX_ERR_BUSY;
}

int fhdx_server_stats(fhdx_server* s, uint64_t* steps, uint64_t* cell_updates, uint64_t* bytes_used)
{
    if (!s) return FHDX_ERR_ARG;
    if (steps) *steps = s->steps_done.load();
    if (cell_updates) *cell_updates = s->cell_updates.load();
    if (bytes_used) *bytes_used = s->bytes_used.load();
    return FHDX_OK;
}

int fhdx_server_destroy(fhdx_server* s)
{
    if (!s) return FHDX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    int rc = engine_destroy(s->e);
    delete s;
    return rc;
}

}  // extern "C"
