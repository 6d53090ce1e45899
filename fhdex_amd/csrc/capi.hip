// capi.hip -- the C ABI of include/fhdx.h.
//
// Host C++ over the HIP kernels: argument and shape checks, the RK3 stage driver of
// timeStep.cpp:11-427, context-owned workspace (allocated once, never per step) and the
// device error flag that replaces the reference's amrex::Abort calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "ctx.hpp"

namespace fhdx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace fhdx

using namespace fhdx;

extern "C" {

int fhdx_abi_version(void) { return FHDX_ABI_VERSION; }

const char* fhdx_status_string(int s)
{
    switch (s) {
        case FHDX_OK: return "ok";
        case FHDX_ERR_ARG: return "invalid argument or array shape";
        case FHDX_ERR_UNSUPPORTED: return "parameters outside the implemented scope";
        case FHDX_ERR_SINGULAR: return "singular matrix in the Valk-Waldmann transport solves";
        case FHDX_ERR_HIP: return "HIP runtime error";
        case FHDX_ERR_MASSFRAC: return "mass fraction outside (0,1) in the transport step";
        case FHDX_ERR_NOMEM: return "device memory budget exhausted";
        case FHDX_ERR_NONFINITE: return "non-finite state";
        case FHDX_ERR_BUSY: return "no free slot";
        default: return "unknown status";
    }
}

const char* fhdx_last_error(void) { return g_last_error.c_str(); }

void fhdx_params_default(fhdx_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->nspecies = 2;
    for (int d = 0; d < 3; ++d) {
        p->n_cells[d] = 32;
        p->dx[d] = 8e-6;
        p->bc_vel_lo[d] = p->bc_vel_hi[d] = -1;
        p->bc_mass_lo[d] = p->bc_mass_hi[d] = -1;
        p->bc_therm_lo[d] = p->bc_therm_hi[d] = -1;
    }
    p->ngc = 2;
    p->algorithm_type = 2;
    p->advection_type = 2;
    p->visc_type = 2;
    p->transport_type = 1;
    p->stoch_stress_form = 1;
    for (int n = 0; n < FHDX_MAX_SPECIES; ++n) {
        p->molmass[n] = 1.;
        p->diameter[n] = 1.;
        p->dof[n] = 3.;
        p->hcv[n] = -1.;
        p->hcp[n] = -1.;
        p->e0[n] = 0.;
    }
    p->k_B = 1.38064852e-16;
    p->Runiv = 8.314462175e7;
    p->variance_coef_mom = p->variance_coef_ener = p->variance_coef_mass = 1.;
    p->zeta_ratio = -1.;
}

void fhdx_params_finalize(fhdx_params* p)
{
    for (int i = 0; i < p->nspecies && i < FHDX_MAX_SPECIES; ++i) {
        if (p->hcv[i] < 0.) p->hcv[i] = 0.5 * p->dof[i] * p->Runiv / p->molmass[i];
        if (p->hcp[i] < 0.) p->hcp[i] = 0.5 * (2. + p->dof[i]) * p->Runiv / p->molmass[i];
    }
    // SetupBC (boundary.cpp:3-12): a periodic direction has periodic mass and thermal codes
    for (int d = 0; d < 3; ++d)
        if (p->bc_vel_lo[d] == -1) {
            p->bc_mass_lo[d] = p->bc_mass_hi[d] = -1;
            p->bc_therm_lo[d] = p->bc_therm_hi[d] = -1;
        }
    // SetupCWall (boundary.cpp:14-170): the wall composition of bc_mass 2 faces from whichever
    // of X or Y sums to 1 (X first); GetMassfrac / GetMolfrac, compressible_functions.H:139-167
    const int NS = p->nspecies < FHDX_MAX_SPECIES ? p->nspecies : FHDX_MAX_SPECIES;
    for (int f = 0; f < 6; ++f) {
        const int code = (f & 1) ? p->bc_mass_hi[f >> 1] : p->bc_mass_lo[f >> 1];
        if (code != 2 && code != 3) continue;
        double sumx = 0., sumy = 0.;
        for (int n = 0; n < NS; ++n) {
            sumx = sumx + p->bc_Xk[f][n];
            sumy = sumy + p->bc_Yk[f][n];
        }
        if (std::fabs(sumx - 1) < 1.e-10) {
            double molmix = 0.;
            for (int n = 0; n < NS; ++n) molmix += p->bc_Xk[f][n] * p->molmass[n];
            for (int n = 0; n < NS; ++n) p->bc_Yk[f][n] = p->bc_Xk[f][n] * p->molmass[n] / molmix;
        } else if (std::fabs(sumy - 1) < 1.e-10) {
            double molmix = 0.;
            for (int n = 0; n < NS; ++n) molmix += p->bc_Yk[f][n] / p->molmass[n];
            molmix = 1. / molmix;
            for (int n = 0; n < NS; ++n) p->bc_Xk[f][n] = p->bc_Yk[f][n] * molmix / p->molmass[n];
        }
        if (code != 3) continue;
        // reservoir gas: density from p and T (GetDensity), or T from p and rho (boundary.cpp:49-57)
        const int d = f >> 1;
        double& pw = (f & 1) ? p->p_hi[d] : p->p_lo[d];
        double& rw = (f & 1) ? p->rho_hi[d] : p->rho_lo[d];
        double& tw = (f & 1) ? p->t_hi[d] : p->t_lo[d];
        double molmix = 0.;
        for (int n = 0; n < NS; ++n) molmix += p->bc_Yk[f][n] / p->molmass[n];
        molmix = 1. / molmix;
        if (pw > 0. && rw <= 0. && tw > 0.) rw = pw / (p->Runiv / molmix) / tw;
        else if (pw > 0. && rw > 0. && tw <= 0.) tw = pw * (molmix / p->Runiv) / rw;
    }
}

int fhdx_params_check(const fhdx_params* p)
{
    if (!p) return FHDX_ERR_ARG;
    if (p->nspecies < 2 || p->nspecies > FHDX_MAX_SPECIES) {
        set_error("nspecies must be 2..8 (a single gas is written as two identical species)");
        return FHDX_ERR_UNSUPPORTED;
    }
    for (int d = 0; d < 3; ++d) {
        if (p->n_cells[d] < 1) {
            set_error("n_cells must be >= 1 in every direction");
            return FHDX_ERR_ARG;
        }
        if (!(p->dx[d] > 0.)) {
            set_error("dx must be positive");
            return FHDX_ERR_ARG;
        }
        const bool per_lo = p->bc_vel_lo[d] == -1, per_hi = p->bc_vel_hi[d] == -1;
        if (per_lo != per_hi) {
            set_error("a direction is periodic on both sides or on neither (bc_vel_lo/hi = -1 together)");
            return FHDX_ERR_ARG;
        }
        if (per_lo) {
            if (p->bc_mass_lo[d] != -1 || p->bc_mass_hi[d] != -1 || p->bc_therm_lo[d] != -1 ||
                p->bc_therm_hi[d] != -1) {
                set_error("a periodic direction needs bc_mass and bc_therm -1 (fhdx_params_finalize sets them)");
                return FHDX_ERR_ARG;
            }
            continue;
        }
        for (int side = 0; side < 2; ++side) {
            const int v = side ? p->bc_vel_hi[d] : p->bc_vel_lo[d];
            const int m = side ? p->bc_mass_hi[d] : p->bc_mass_lo[d];
            const int t = side ? p->bc_therm_hi[d] : p->bc_therm_lo[d];
            if (v != 1 && v != 2) {
                set_error("bc_vel of a wall must be 1 (slip) or 2 (no slip)");
                return FHDX_ERR_UNSUPPORTED;
            }
            if (m == 3 && d != 0) {
                set_error("reservoirs (bc_mass 3) are implemented on x faces only: the reference's y and z "
                          "reservoir ghosts keep stale velocity and energy (boundary.cpp:648-649, 806-807)");
                return FHDX_ERR_UNSUPPORTED;
            }
            if (m != 1 && m != 2 && m != 3) {
                set_error("bc_mass of a wall must be 1 (wall), 2 (fixed composition) or 3 (x reservoir)");
                return FHDX_ERR_UNSUPPORTED;
            }
            if (m == 3) {
                const double pw = side ? p->p_hi[d] : p->p_lo[d], rw = side ? p->rho_hi[d] : p->rho_lo[d];
                const double tr = side ? p->t_hi[d] : p->t_lo[d];
                if (!(pw > 0.) || !(rw > 0.) || !(tr > 0.)) {
                    set_error("a reservoir needs p_lo/p_hi > 0 and, after fhdx_params_finalize, rho and t > 0");
                    return FHDX_ERR_ARG;
                }
            }
            if (t != 1 && t != 2) {
                set_error("bc_therm of a wall must be 1 (adiabatic) or 2 (isothermal)");
                return FHDX_ERR_UNSUPPORTED;
            }
            const double tw = side ? p->t_hi[d] : p->t_lo[d];
            if (t == 2 && !(tw > 0.)) {
                set_error("an isothermal wall needs t_lo/t_hi > 0");
                return FHDX_ERR_ARG;
            }
            if (m == 2 || m == 3) {
                const int f = 2 * d + side;
                double sy = 0.;
                bool pos = true;
                for (int n = 0; n < p->nspecies && n < FHDX_MAX_SPECIES; ++n) {
                    sy += p->bc_Yk[f][n];
                    pos = pos && p->bc_Yk[f][n] > 0. && p->bc_Xk[f][n] > 0.;
                }
                if (!pos || std::fabs(sy - 1.) > 1e-10) {
                    set_error("a bc_mass 2/3 face needs positive bc_Yk/bc_Xk summing to 1 (fhdx_params_finalize)");
                    return FHDX_ERR_ARG;
                }
            }
        }
        if (p->algorithm_type != 2) {
            set_error("walls need algorithm_type 2 (setBC writes the species ghosts only then)");
            return FHDX_ERR_UNSUPPORTED;
        }
    }
    if (p->ngc < 2) {
        set_error("ngc must be >= 2 (the stencils reach two cells)");
        return FHDX_ERR_UNSUPPORTED;
    }
    for (int d = 0; d < 3; ++d)
        if (p->bc_vel_lo[d] != -1 && 2 * p->ngc > p->n_cells[d]) {
            set_error("a direction with walls needs at least 2 ngc cells (setBC mirrors ngc valid cells at each wall)");
            return FHDX_ERR_UNSUPPORTED;
        }
    {
        const int g = p->ngc;
        const double plane = 8.0 * (p->n_cells[0] + 2.0 * g) * (p->n_cells[1] + 2.0 * g) * (p->n_cells[2] + 2.0 * g);
        if (plane >= 4294967296.0) {
            set_error("one component of a cell array must stay below 4 GiB (box too large for one context)");
            return FHDX_ERR_UNSUPPORTED;
        }
    }
    if (p->transport_type < 1 || p->transport_type > 3) {
        set_error("transport_type must be 1 (Giovangigli), 2 (Valk-Waldmann) or 3 (Hirschfelder-Curtiss-Bird, "
                  "binary)");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (p->transport_type == 3 && p->nspecies != 2) {
        set_error("transport_type 3 (Hirschfelder-Curtiss-Bird) is for binary mixtures (nspecies 2)");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (p->nreaction < 0 || p->nreaction > FHDX_MAX_REACTION) {
        set_error("nreaction must be 0..FHDX_MAX_REACTION");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (p->advection_type < 0 || p->advection_type > 2) return FHDX_ERR_ARG;
    if (p->stoch_stress_form != 0 && p->stoch_stress_form != 1) return FHDX_ERR_ARG;
    for (int n = 0; n < p->nspecies; ++n)
        if (!(p->hcv[n] > 0.) || !(p->hcp[n] > 0.) || !(p->molmass[n] > 0.) || !(p->diameter[n] > 0.)) {
            set_error("species properties must be positive (call fhdx_params_finalize for hcv/hcp)");
            return FHDX_ERR_ARG;
        }
    return FHDX_OK;
}

}  // extern "C"

// ---- shape checks ----------------------------------------------------------------------

namespace {

// A valid box of the domain (global cell indices): lower corner and cells per direction.  A
// MultiFab split into several boxes hands each FAB to the library on its own; every array of
// one call belongs to the same box.
struct Box {
    int lo[3], n[3];
};

Box domain_box(const fhdx_ctx* c)
{
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = 0;
        b.n[d] = c->p.n_cells[d];
    }
    return b;
}

bool is_domain(const fhdx_ctx* c, const Box& b)
{
    for (int d = 0; d < 3; ++d)
        if (b.lo[d] != 0 || b.n[d] != c->p.n_cells[d]) return false;
    return true;
}

// A box must lie inside the domain, and a box touching a wall must hold the ngc cells that setBC
// mirrors into the ghosts there (boundary.cpp:172-1209).
bool box_ok(const fhdx_ctx* c, const Box& b, const char* what)
{
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
        ok = ok && b.n[d] >= 1 && b.lo[d] >= 0 && b.lo[d] + b.n[d] <= c->p.n_cells[d];
        const bool at_wall = !c->dp.periodic[d] && (b.lo[d] == 0 || b.lo[d] + b.n[d] == c->p.n_cells[d]);
        ok = ok && !(at_wall && b.n[d] < c->p.ngc);
    }
    if (!ok)
        set_error(std::string("bad box: ") + what +
                  " (a box must lie inside the domain and hold >= ngc cells across a wall it touches)");
    return ok;
}

// The valid box of a cell FAB with ngc ghost layers.
bool cell_box(const fhdx_ctx* c, const fhdx_fab& f, int ncomp, const char* what, Box& b)
{
    const int g = c->p.ngc;
    if (!f.p || f.ncomp < ncomp) {
        set_error(std::string("bad cell array: ") + what);
        return false;
    }
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = f.lo[d] + g;
        b.n[d] = f.hi[d] - f.lo[d] + 1 - 2 * g;
    }
    return box_ok(c, b, what);
}

// The valid box of a node FAB (no ghosts).
bool node_box(const fhdx_ctx* c, const fhdx_fab& f, const char* what, Box& b)
{
    if (!f.p || f.ncomp < 1) {
        set_error(std::string("bad node array: ") + what);
        return false;
    }
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = f.lo[d];
        b.n[d] = f.hi[d] - f.lo[d];
    }
    return box_ok(c, b, what);
}

bool check_cell(const fhdx_ctx* c, const Box& b, const fhdx_fab& f, int ncomp, const char* what)
{
    const int g = c->p.ngc;
    bool ok = f.p != nullptr && f.ncomp >= ncomp;
    for (int d = 0; d < 3; ++d) ok = ok && f.lo[d] == b.lo[d] - g && f.hi[d] == b.lo[d] + b.n[d] - 1 + g;
    if (!ok) set_error(std::string("bad cell array shape: ") + what);
    return ok;
}

bool check_face(const fhdx_ctx* c, const Box& b, const fhdx_fab& f, int dir, int ncomp, const char* what)
{
    bool ok = f.p != nullptr && f.ncomp >= ncomp;
    for (int d = 0; d < 3; ++d) ok = ok && f.lo[d] == b.lo[d] && f.hi[d] == b.lo[d] + b.n[d] - 1 + (d == dir);
    if (!ok) set_error(std::string("bad face array shape: ") + what);
    return ok;
}

bool check_node(const fhdx_ctx* c, const Box& b, const fhdx_fab& f, const char* what)
{
    bool ok = f.p != nullptr && f.ncomp >= 1;
    for (int d = 0; d < 3; ++d) ok = ok && f.lo[d] == b.lo[d] && f.hi[d] == b.lo[d] + b.n[d];
    if (!ok) set_error(std::string("bad node array shape: ") + what);
    return ok;
}

// Whole-domain forms (entry points that fill periodic ghosts themselves take one box only).
bool check_cell(const fhdx_ctx* c, const fhdx_fab& f, int ncomp, const char* what)
{
    return check_cell(c, domain_box(c), f, ncomp, what);
}

bool check_face(const fhdx_ctx* c, const fhdx_fab& f, int dir, int ncomp, const char* what)
{
    return check_face(c, domain_box(c), f, dir, ncomp, what);
}

bool check_node(const fhdx_ctx* c, const fhdx_fab& f, const char* what)
{
    return check_node(c, domain_box(c), f, what);
}

// The context's device parameters with the launch box set to b.
DevParams box_params(const fhdx_ctx* c, const Box& b)
{
    DevParams P = c->dp;
    for (int d = 0; d < 3; ++d) {
        P.bo[d] = b.lo[d];
        P.bn[d] = b.n[d];
    }
    return P;
}

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int alloc_fab(std::vector<double*>& blocks, fhdx_fab& f, int l0, int l1, int l2, int h0, int h1, int h2, int nc)
{
    f.lo[0] = l0; f.lo[1] = l1; f.lo[2] = l2;
    f.hi[0] = h0; f.hi[1] = h1; f.hi[2] = h2;
    f.ncomp = nc;
    f.pad = 0;
    size_t n = (size_t)(h0 - l0 + 1) * (h1 - l1 + 1) * (h2 - l2 + 1) * nc;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(double));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc of workspace failed: ") + hipGetErrorString(e));
        return FHDX_ERR_NOMEM;
    }
    (void)hipMemset(p, 0, n * sizeof(double));
    f.p = (double*)p;
    blocks.push_back((double*)p);
    return FHDX_OK;
}

int ensure_ref_workspace(fhdx_ctx* c)
{
    RefWorkspace& w = c->ref;
    if (w.ready) return FHDX_OK;
    const int g = c->p.ngc, NS = c->p.nspecies, nv = 5 + NS;
    const int n0 = c->p.n_cells[0], n1 = c->p.n_cells[1], n2 = c->p.n_cells[2];
    int r = 0;
    auto cell = [&](fhdx_fab& f, int nc) {
        if (!r) r = alloc_fab(w.blocks, f, -g, -g, -g, n0 - 1 + g, n1 - 1 + g, n2 - 1 + g, nc);
    };
    auto face = [&](fhdx_fab& f, int d, int nc) {
        if (!r) r = alloc_fab(w.blocks, f, 0, 0, 0, n0 - 1 + (d == 0), n1 - 1 + (d == 1), n2 - 1 + (d == 2), nc);
    };
    auto node = [&](fhdx_fab& f) {
        if (!r) r = alloc_fab(w.blocks, f, 0, 0, 0, n0, n1, n2, 1);
    };
    for (int d = 0; d < 3; ++d) {
        face(w.sA[d], d, nv);
        face(w.sB[d], d, nv);
    }
    node(w.rA);
    node(w.rB);
    if (c->p.nreaction > 0) {
        cell(w.source, nv);  // components 0..4 stay zero (main_driver.cpp:241-242)
        cell(w.ranA, c->p.nreaction);
        cell(w.ranB, c->p.nreaction);
        cell(w.ran, c->p.nreaction);
    }
    if (r) return r;
    FHDX_HIP_TRY(hipDeviceSynchronize());
    w.ready = true;
    return FHDX_OK;
}

void free_ref_workspace(RefWorkspace& w)
{
    for (double* p : w.blocks) (void)hipFree(p);
    w.blocks.clear();
    w.ready = false;
}

long npts(const fhdx_fab& f)
{
    return (long)(f.hi[0] - f.lo[0] + 1) * (f.hi[1] - f.lo[1] + 1) * (f.hi[2] - f.lo[2] + 1);
}

// The arrays of one reference-shaped RK3 stage on one box: the reference's own argument list
// of RK3step (compressible_functions.H:41-51), the step's noise sets (stochFlux_A/B,
// rancorn_A/B of timeStep.cpp:38-53) and, with chemistry, source, ranchem and its noise sets.
struct StageArrays {
    fhdx_fab cu, cup, cup2, prim, eta, zeta, kappa, chi, D;
    fhdx_fab flux[3], stochFlux[3], cornx[3], corny[3], cornz[3], visccorn, rancorn;
    fhdx_fab sA[3], sB[3], rA, rB;
    fhdx_fab source, ranchem, ranchemA, ranchemB;
};

// Stage st (0, 1, 2) of RK3step on the launch box of P, from the transport coefficients of the
// stage input to conservedToPrimitive of its output (timeStep.cpp:99-202, 211-307, 316-415):
// transport on the grown box, the stage-weighted noise, calculateFlux, the chemistry source,
// the stage combination and prim on the valid cells.  The ghost cells of the output state and
// of prim are left to the caller (FillBoundary, then setBC: timeStep.cpp:204-209, 309-314,
// 417-425).
int rk3_stage(const fhdx_ctx* c, const DevParams& P, const StageArrays& A, int st, double dt, hipStream_t s)
{
    const bool chem = c->p.nreaction > 0;
    const int nv = P.nvars;
    double ws[3];
    stage_weights(ws);
    const fhdx_fab in[3] = {A.cu, A.cup, A.cup2};
    const fhdx_fab out[3] = {A.cup, A.cup2, A.cu};
    FluxArgs a;
    a.prim = make_arr(A.prim);
    a.eta = make_arr(A.eta);
    a.zeta = make_arr(A.zeta);
    a.kappa = make_arr(A.kappa);
    a.chi = make_arr(A.chi);
    a.D = make_arr(A.D);
    Arr fl[3];
    for (int d = 0; d < 3; ++d) {
        a.flux[d] = fl[d] = make_arr(A.flux[d]);
        a.ran[d] = make_arr(A.stochFlux[d]);
        a.corn[0][d] = make_arr(A.cornx[d]);
        a.corn[1][d] = make_arr(A.corny[d]);
        a.corn[2][d] = make_arr(A.cornz[d]);
    }
    a.visccorn = make_arr(A.visccorn);
    a.rancorn = make_arr(A.rancorn);
    a.dt = dt;
    a.cons = make_arr(in[st]);
    int r;
    if ((r = ref_transport(P, a.prim, a.eta, a.zeta, a.kappa, a.chi, a.D, c->d_err, s))) return r;
    if (P.stoch_on) {
        // the stochastic and node passes form stochFlux = A + w B and rancorn themselves
        a.weigh = true;
        for (int d = 0; d < 3; ++d) {
            a.nA[d] = make_arr(A.sA[d]);
            a.nB[d] = make_arr(A.sB[d]);
        }
        a.rA = make_arr(A.rA);
        a.rB = make_arr(A.rB);
        a.w = ws[st];
    } else {
        double* const dst[4] = {A.stochFlux[0].p, A.stochFlux[1].p, A.stochFlux[2].p, A.rancorn.p};
        const double* const la[4] = {A.sA[0].p, A.sA[1].p, A.sA[2].p, A.rA.p};
        const double* const lb[4] = {A.sB[0].p, A.sB[1].p, A.sB[2].p, A.rB.p};
        const long np[4] = {npts(A.stochFlux[0]), npts(A.stochFlux[1]), npts(A.stochFlux[2]), npts(A.rancorn)};
        const int lnc[4] = {nv, nv, nv, 1}, lc0[4] = {1, 1, 1, 0};
        if ((r = ref_lincomb4(dst, la, lb, np, lnc, lc0, ws[st], s))) return r;
    }
    const int fnc[3] = {A.flux[0].ncomp, A.flux[1].ncomp, A.flux[2].ncomp};
    if ((r = ref_calculate_flux(P, a, const_cast<int*>(fnc), s))) return r;
    Arr src{};
    if (chem) {  // timeStep.cpp:135-143: ranchem = A + w B, then the source from this stage's prim
        src = make_arr(A.source);
        if ((r = ref_lincomb(A.ranchem.p, A.ranchemA.p, A.ranchemB.p, npts(A.ranchem), c->p.nreaction, 0, ws[st],
                             s)))
            return r;
        if ((r = ref_chem_source(P, a.prim, src, make_arr(A.ranchem), dt, s))) return r;
    }
    if ((r = ref_stage(P, make_arr(out[st]), make_arr(A.cu), make_arr(in[st]), fl, src, dt, st, s))) return r;
    return ref_c2p(P, a.prim, make_arr(out[st]), s);
}

// The reference RK3step with the reference's arrays (timeStep.cpp:11-427) on a single-box
// domain: each stage, then the periodic images and setBC of the new state and prim.
int rk3_ref(fhdx_ctx* c, const StageArrays& A, double dt, hipStream_t s)
{
    const DevParams& P = c->dp;
    const fhdx_fab out[3] = {A.cup, A.cup2, A.cu};
    for (int st = 0; st < 3; ++st) {
        int r;
        if ((r = rk3_stage(c, P, A, st, dt, s))) return r;
        if ((r = ref_fill_boundary2(P, make_arr(out[st]), P.nvars, make_arr(A.prim), P.nprim, s))) return r;
        if ((r = ref_set_bc(P, make_arr(A.prim), make_arr(out[st]), s))) return r;  // timeStep.cpp:209, 314, 425
    }
    return FHDX_OK;
}

// The fused step, the ensemble engine and the server serve periodic boxes of at least
// max(4, 2 ngc) cells per direction: their producing kernels write each ghost cell's single
// periodic image.  Walls and thinner (quasi-2-D) boxes take the reference-shaped step.
int reject_unfused(const fhdx_ctx* c, const char* who)
{
    if (c->dp.any_wall) {
        set_error(std::string(who) + ": walls are served by the reference-shaped step (fhdx_rk3_step_ref)");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (!fused_box_ok(c->p)) {
        set_error(std::string(who) + ": boxes thinner than max(4, 2 ngc) cells in a direction are served by "
                                     "the reference-shaped step (fhdx_rk3_step_ref)");
        return FHDX_ERR_UNSUPPORTED;
    }
    return FHDX_OK;
}

}  // namespace

extern "C" {

int fhdx_ctx_create(const fhdx_params* p, fhdx_ctx** out)
{
    if (!out) return FHDX_ERR_ARG;
    *out = nullptr;
    int r = fhdx_params_check(p);
    if (r) return r;
    fhdx_ctx* c = new fhdx_ctx();
    c->p = *p;
    c->dp = make_dev_params(*p);
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_err, sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->d_err, 0, sizeof(int));
    if (e != hipSuccess) {
        set_error(std::string("context setup failed: ") + hipGetErrorString(e));
        if (c->d_err) (void)hipFree(c->d_err);
        delete c;
        return FHDX_ERR_HIP;
    }
    // this device's copy of the no-allocation scratch, looked up once per device
    if ((r = noalloc_scratch_init(c->device))) {
        (void)hipFree(c->d_err);
        delete c;
        return r;
    }
    *out = c;
    return FHDX_OK;
}

int fhdx_ctx_destroy(fhdx_ctx* c)
{
    if (!c) return FHDX_OK;
    (void)hipDeviceSynchronize();
    free_ref_workspace(c->ref);
    fused_free(c->fused);
    stag_workspace_free(c->stag);
    if (c->st_yz) (void)hipFree(c->st_yz);
    if (c->st_part) (void)hipFree(c->st_part);
    if (c->st_misc) (void)hipFree(c->st_misc);
    if (c->d_err) (void)hipFree(c->d_err);
    delete c;
    return FHDX_OK;
}

int fhdx_ctx_check(fhdx_ctx* c, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    FHDX_HIP_TRY(hipStreamSynchronize(S(stream)));
    int h = 0;
    FHDX_HIP_TRY(hipMemcpy(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (h) {
        FHDX_HIP_TRY(hipMemset(c->d_err, 0, sizeof(int)));
        if (h & 1) {
            set_error("mass fraction outside (0,1) met in the transport step");
            return FHDX_ERR_MASSFRAC;
        }
        if (h & 2) {
            set_error("singular matrix in the Valk-Waldmann transport solves (transport_type 2)");
            return FHDX_ERR_SINGULAR;
        }
        return FHDX_ERR_NONFINITE;
    }
    return FHDX_OK;
}

int fhdx_cons_to_prim(fhdx_ctx* c, fhdx_fab prim, fhdx_fab cons, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    Box b;
    if (!cell_box(c, prim, c->dp.nprim, "prim", b) || !check_cell(c, b, cons, c->dp.nvars, "cons"))
        return FHDX_ERR_ARG;
    return ref_c2p(box_params(c, b), make_arr(prim), make_arr(cons), S(stream));
}

int fhdx_transport(fhdx_ctx* c, fhdx_fab prim, fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi,
                   fhdx_fab D, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    const int NS = c->p.nspecies;
    Box b;
    if (!cell_box(c, prim, c->dp.nprim, "prim", b) || !check_cell(c, b, eta, 1, "eta") ||
        !check_cell(c, b, zeta, 1, "zeta") || !check_cell(c, b, kappa, 1, "kappa") ||
        !check_cell(c, b, chi, NS, "chi") || !check_cell(c, b, D, NS * NS, "D"))
        return FHDX_ERR_ARG;
    return ref_transport(box_params(c, b), make_arr(prim), make_arr(eta), make_arr(zeta), make_arr(kappa), make_arr(chi),
                         make_arr(D), c->d_err, S(stream));
}

int fhdx_fill_boundary(fhdx_ctx* c, fhdx_fab mf, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    if (!check_cell(c, mf, 1, "mf (the whole domain: a box of a split domain takes its ghosts from the others)"))
        return FHDX_ERR_ARG;
    return ref_fill_boundary(c->dp, make_arr(mf), mf.ncomp, S(stream));
}

int fhdx_set_bc(fhdx_ctx* c, fhdx_fab prim, fhdx_fab cons, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    Box b;
    if (!cell_box(c, prim, c->dp.nprim, "prim", b) || !check_cell(c, b, cons, c->dp.nvars, "cons"))
        return FHDX_ERR_ARG;
    return ref_set_bc(box_params(c, b), make_arr(prim), make_arr(cons), S(stream));
}

int fhdx_calculate_flux(fhdx_ctx* c, fhdx_fab cons, fhdx_fab prim, fhdx_fab eta, fhdx_fab zeta,
                        fhdx_fab kappa, fhdx_fab chi, fhdx_fab D, const fhdx_fab flux[3],
                        const fhdx_fab stochFlux[3], const fhdx_fab cornx[3], const fhdx_fab corny[3],
                        const fhdx_fab cornz[3], fhdx_fab visccorn, fhdx_fab rancorn, double dt, void* stream)
{
    if (!c || !flux || !stochFlux || !cornx || !corny || !cornz) return FHDX_ERR_ARG;
    const int NS = c->p.nspecies, nv = c->dp.nvars;
    Box b;
    bool ok = cell_box(c, cons, nv, "cons", b) && check_cell(c, b, prim, c->dp.nprim, "prim") &&
              check_cell(c, b, eta, 1, "eta") && check_cell(c, b, zeta, 1, "zeta") &&
              check_cell(c, b, kappa, 1, "kappa") && check_cell(c, b, chi, NS, "chi") &&
              check_cell(c, b, D, NS * NS, "D") && check_node(c, b, visccorn, "visccorn") &&
              check_node(c, b, rancorn, "rancorn");
    for (int d = 0; ok && d < 3; ++d)
        ok = check_face(c, b, flux[d], d, nv + 4, "flux") && check_face(c, b, stochFlux[d], d, nv, "stochFlux") &&
             check_node(c, b, cornx[d], "cornx") && check_node(c, b, corny[d], "corny") &&
             check_node(c, b, cornz[d], "cornz");
    if (!ok) return FHDX_ERR_ARG;
    if (!(dt > 0.)) return FHDX_ERR_ARG;
    FluxArgs a;
    a.cons = make_arr(cons);
    a.prim = make_arr(prim);
    a.eta = make_arr(eta);
    a.zeta = make_arr(zeta);
    a.kappa = make_arr(kappa);
    a.chi = make_arr(chi);
    a.D = make_arr(D);
    for (int d = 0; d < 3; ++d) {
        a.flux[d] = make_arr(flux[d]);
        a.ran[d] = make_arr(stochFlux[d]);
        a.corn[0][d] = make_arr(cornx[d]);
        a.corn[1][d] = make_arr(corny[d]);
        a.corn[2][d] = make_arr(cornz[d]);
    }
    a.visccorn = make_arr(visccorn);
    a.rancorn = make_arr(rancorn);
    a.dt = dt;
    int fnc[3] = {flux[0].ncomp, flux[1].ncomp, flux[2].ncomp};
    return ref_calculate_flux(box_params(c, b), a, fnc, S(stream));
}

int fhdx_fill_noise(fhdx_ctx* c, uint64_t seed, uint64_t step, uint32_t instance, const fhdx_fab stochA[3],
                    const fhdx_fab stochB[3], fhdx_fab rancornA, fhdx_fab rancornB, void* stream)
{
    if (!c || !stochA || !stochB) return FHDX_ERR_ARG;
    const int nv = c->dp.nvars;
    Box b;
    bool ok = node_box(c, rancornA, "rancornA", b) && check_node(c, b, rancornB, "rancornB");
    for (int d = 0; ok && d < 3; ++d)
        ok = check_face(c, b, stochA[d], d, nv, "stochA") && check_face(c, b, stochB[d], d, nv, "stochB");
    if (!ok) return FHDX_ERR_ARG;
    hipStream_t s = S(stream);
    Arr A[3], B[3];
    for (int d = 0; d < 3; ++d) {
        A[d] = make_arr(stochA[d]);
        B[d] = make_arr(stochB[d]);
        // the noise kernel writes components 0..nvars-1; zero any further ones first
        int r = FHDX_OK;
        if (stochA[d].ncomp > c->dp.nvars) r = fill_value(stochA[d].p, A[d].sn * stochA[d].ncomp, 0.0, s);
        if (!r && stochB[d].ncomp > c->dp.nvars) r = fill_value(stochB[d].p, B[d].sn * stochB[d].ncomp, 0.0, s);
        if (r) return r;
    }
    return ref_fill_noise(box_params(c, b), make_noise_key(seed, step, instance), A, B, make_arr(rancornA),
                          make_arr(rancornB), s);
}

int fhdx_rk3_step_ref(fhdx_ctx* c, fhdx_fab cu, fhdx_fab cup, fhdx_fab cup2, fhdx_fab prim, fhdx_fab eta,
                      fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D, const fhdx_fab flux[3],
                      const fhdx_fab stochFlux[3], const fhdx_fab cornx[3], const fhdx_fab corny[3],
                      const fhdx_fab cornz[3], fhdx_fab visccorn, fhdx_fab rancorn, const fhdx_fab* stochA,
                      const fhdx_fab* stochB, const fhdx_fab* rancornA, const fhdx_fab* rancornB, double dt,
                      uint64_t seed, uint64_t step, void* stream)
{
    if (!c || !flux || !stochFlux || !cornx || !corny || !cornz) return FHDX_ERR_ARG;
    const int NS = c->p.nspecies, nv = c->dp.nvars;
    bool ok = check_cell(c, cu, nv, "cu") && check_cell(c, cup, nv, "cup") && check_cell(c, cup2, nv, "cup2") &&
              check_cell(c, prim, c->dp.nprim, "prim") && check_cell(c, eta, 1, "eta") &&
              check_cell(c, zeta, 1, "zeta") && check_cell(c, kappa, 1, "kappa") && check_cell(c, chi, NS, "chi") &&
              check_cell(c, D, NS * NS, "D") && check_node(c, visccorn, "visccorn") &&
              check_node(c, rancorn, "rancorn");
    for (int d = 0; ok && d < 3; ++d)
        ok = check_face(c, flux[d], d, nv + 4, "flux") && check_face(c, stochFlux[d], d, nv, "stochFlux") &&
             check_node(c, cornx[d], "cornx") && check_node(c, corny[d], "corny") && check_node(c, cornz[d], "cornz");
    if (!ok || !(dt > 0.)) return FHDX_ERR_ARG;
    hipStream_t s = S(stream);
    const bool injected = stochA != nullptr;
    if (c->p.nreaction > 0 && !c->chem_set) {
        set_error("nreaction > 0: call fhdx_ctx_set_chemistry before stepping");
        return FHDX_ERR_ARG;
    }
    StageArrays A{};
    A.cu = cu;
    A.cup = cup;
    A.cup2 = cup2;
    A.prim = prim;
    A.eta = eta;
    A.zeta = zeta;
    A.kappa = kappa;
    A.chi = chi;
    A.D = D;
    for (int d = 0; d < 3; ++d) {
        A.flux[d] = flux[d];
        A.stochFlux[d] = stochFlux[d];
        A.cornx[d] = cornx[d];
        A.corny[d] = corny[d];
        A.cornz[d] = cornz[d];
    }
    A.visccorn = visccorn;
    A.rancorn = rancorn;
    if (injected) {
        if (c->p.nreaction > 0) {
            set_error("caller noise with chemistry: the reaction noise has no array argument");
            return FHDX_ERR_UNSUPPORTED;
        }
        if (!stochB || !rancornA || !rancornB) return FHDX_ERR_ARG;
        for (int d = 0; d < 3; ++d)
            if (!check_face(c, stochA[d], d, nv, "stochA") || !check_face(c, stochB[d], d, nv, "stochB"))
                return FHDX_ERR_ARG;
        if (!check_node(c, *rancornA, "rancornA") || !check_node(c, *rancornB, "rancornB")) return FHDX_ERR_ARG;
        for (int d = 0; d < 3; ++d) {
            A.sA[d] = stochA[d];
            A.sB[d] = stochB[d];
        }
        A.rA = *rancornA;
        A.rB = *rancornB;
        return rk3_ref(c, A, dt, s);
    }
    int r = ensure_ref_workspace(c);
    if (r) return r;
    RefWorkspace& w = c->ref;
    if ((r = fhdx_fill_noise(c, seed, step, 0, w.sA, w.sB, w.rA, w.rB, stream))) return r;
    if (c->p.nreaction > 0 &&
        (r = ref_fill_chem_noise(c->dp, make_noise_key(seed, step, 0), make_arr(w.ranA), make_arr(w.ranB), s)))
        return r;
    for (int d = 0; d < 3; ++d) {
        A.sA[d] = w.sA[d];
        A.sB[d] = w.sB[d];
    }
    A.rA = w.rA;
    A.rB = w.rB;
    A.source = w.source;
    A.ranchem = w.ran;
    A.ranchemA = w.ranA;
    A.ranchemB = w.ranB;
    return rk3_ref(c, A, dt, s);
}

int fhdx_rk3_stage_ref(fhdx_ctx* c, int stage, fhdx_fab cu, fhdx_fab cup, fhdx_fab cup2, fhdx_fab prim,
                       fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D, const fhdx_fab flux[3],
                       const fhdx_fab stochFlux[3], const fhdx_fab cornx[3], const fhdx_fab corny[3],
                       const fhdx_fab cornz[3], fhdx_fab visccorn, fhdx_fab rancorn, const fhdx_fab stochA[3],
                       const fhdx_fab stochB[3], fhdx_fab rancornA, fhdx_fab rancornB, fhdx_fab source,
                       fhdx_fab ranchem, fhdx_fab ranchemA, fhdx_fab ranchemB, double dt, void* stream)
{
    if (!c || !flux || !stochFlux || !cornx || !corny || !cornz || !stochA || !stochB) return FHDX_ERR_ARG;
    if (stage < 0 || stage > 2 || !(dt > 0.)) {
        set_error("fhdx_rk3_stage_ref: stage must be 0, 1 or 2 and dt > 0");
        return FHDX_ERR_ARG;
    }
    const int NS = c->p.nspecies, nv = c->dp.nvars, nr = c->p.nreaction;
    Box b;
    bool ok = cell_box(c, cu, nv, "cu", b) && check_cell(c, b, cup, nv, "cup") &&
              check_cell(c, b, cup2, nv, "cup2") && check_cell(c, b, prim, c->dp.nprim, "prim") &&
              check_cell(c, b, eta, 1, "eta") && check_cell(c, b, zeta, 1, "zeta") &&
              check_cell(c, b, kappa, 1, "kappa") && check_cell(c, b, chi, NS, "chi") &&
              check_cell(c, b, D, NS * NS, "D") && check_node(c, b, visccorn, "visccorn") &&
              check_node(c, b, rancorn, "rancorn") && check_node(c, b, rancornA, "rancornA") &&
              check_node(c, b, rancornB, "rancornB");
    for (int d = 0; ok && d < 3; ++d)
        ok = check_face(c, b, flux[d], d, nv + 4, "flux") && check_face(c, b, stochFlux[d], d, nv, "stochFlux") &&
             check_face(c, b, stochA[d], d, nv, "stochA") && check_face(c, b, stochB[d], d, nv, "stochB") &&
             check_node(c, b, cornx[d], "cornx") && check_node(c, b, corny[d], "corny") &&
             check_node(c, b, cornz[d], "cornz");
    if (ok && nr > 0) {
        if (!c->chem_set) {
            set_error("nreaction > 0: call fhdx_ctx_set_chemistry before stepping");
            return FHDX_ERR_ARG;
        }
        ok = check_cell(c, b, source, nv, "source") && check_cell(c, b, ranchem, nr, "ranchem") &&
             check_cell(c, b, ranchemA, nr, "ranchemA") && check_cell(c, b, ranchemB, nr, "ranchemB") &&
             ranchemA.ncomp == ranchem.ncomp && ranchemB.ncomp == ranchem.ncomp;
    }
    if (!ok) return FHDX_ERR_ARG;
    StageArrays A{};
    A.cu = cu;
    A.cup = cup;
    A.cup2 = cup2;
    A.prim = prim;
    A.eta = eta;
    A.zeta = zeta;
    A.kappa = kappa;
    A.chi = chi;
    A.D = D;
    for (int d = 0; d < 3; ++d) {
        A.flux[d] = flux[d];
        A.stochFlux[d] = stochFlux[d];
        A.cornx[d] = cornx[d];
        A.corny[d] = corny[d];
        A.cornz[d] = cornz[d];
        A.sA[d] = stochA[d];
        A.sB[d] = stochB[d];
    }
    A.visccorn = visccorn;
    A.rancorn = rancorn;
    A.rA = rancornA;
    A.rB = rancornB;
    if (nr > 0) {
        A.source = source;
        A.ranchem = ranchem;
        A.ranchemA = ranchemA;
        A.ranchemB = ranchemB;
    }
    return rk3_stage(c, box_params(c, b), A, stage, dt, S(stream));
}

int fhdx_rk3_step(fhdx_ctx* c, fhdx_fab cu, fhdx_fab prim, double dt, uint64_t seed, uint64_t step, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    if (!check_cell(c, cu, c->dp.nvars, "cu") || !check_cell(c, prim, c->dp.nprim, "prim") || !(dt > 0.))
        return FHDX_ERR_ARG;
    if (c->p.nreaction > 0 && !c->chem_set) {
        set_error("nreaction > 0: call fhdx_ctx_set_chemistry before stepping");
        return FHDX_ERR_ARG;
    }
    if (int r = reject_unfused(c, "fhdx_rk3_step")) return r;
    return fused_rk3_step(c->p, c->dp, c->fused, cu, prim, dt, make_noise_key(seed, step, 0), nullptr, nullptr,
                          nullptr, nullptr, c->d_err, S(stream));
}

int fhdx_rk3_step_noalloc(fhdx_ctx* c, fhdx_fab cu, fhdx_fab prim, double dt, uint64_t seed, uint64_t step,
                          void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    if (!check_cell(c, cu, c->dp.nvars, "cu") || !check_cell(c, prim, c->dp.nprim, "prim") || !(dt > 0.))
        return FHDX_ERR_ARG;
    if (c->p.nreaction > 0 && !c->chem_set) {
        set_error("nreaction > 0: call fhdx_ctx_set_chemistry before stepping");
        return FHDX_ERR_ARG;
    }
    if (int r = reject_unfused(c, "fhdx_rk3_step_noalloc")) return r;
    int dev = -1;
    FHDX_HIP_TRY(hipGetDevice(&dev));
    if (dev != c->device) {
        set_error("fhdx_rk3_step_noalloc: the current device is not the context's (its scratch lives on that device)");
        return FHDX_ERR_ARG;
    }
    double* base = noalloc_scratch(dev);
    if (!base) {
        set_error("fhdx_rk3_step_noalloc: the scratch of this device was not looked up");
        return FHDX_ERR_HIP;
    }
    return fused_rk3_step_in(c->p, c->dp, c->fused, base, kNoallocScratchBytes, cu, prim, dt,
                             make_noise_key(seed, step, 0), c->d_err, S(stream));
}

int fhdx_rk3_step_injected(fhdx_ctx* c, fhdx_fab cu, fhdx_fab prim, const fhdx_fab stochA[3],
                           const fhdx_fab stochB[3], fhdx_fab rancornA, fhdx_fab rancornB, double dt, void* stream)
{
    if (!c || !stochA || !stochB) return FHDX_ERR_ARG;
    if (c->p.nreaction > 0) {
        set_error("caller noise with chemistry: the reaction noise has no array argument");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (int r = reject_unfused(c, "fhdx_rk3_step_injected")) return r;
    const int nv = c->dp.nvars;
    bool ok = check_cell(c, cu, nv, "cu") && check_cell(c, prim, c->dp.nprim, "prim") &&
              check_node(c, rancornA, "rancornA") && check_node(c, rancornB, "rancornB") && dt > 0.;
    for (int d = 0; ok && d < 3; ++d)
        ok = check_face(c, stochA[d], d, nv, "stochA") && check_face(c, stochB[d], d, nv, "stochB");
    if (!ok) return FHDX_ERR_ARG;
    return fused_rk3_step(c->p, c->dp, c->fused, cu, prim, dt, make_noise_key(0, 0, 0), stochA, stochB, &rancornA,
                          &rancornB, c->d_err, S(stream));
}

// ---- chemistry ------------------------------------------------------------------------------

int fhdx_ctx_set_chemistry(fhdx_ctx* c, const fhdx_chemistry* ch)
{
    if (!c || !ch) return FHDX_ERR_ARG;
    const int nr = ch->nreaction, NS = c->p.nspecies;
    if (nr < 1 || nr > FHDX_MAX_REACTION || nr != c->p.nreaction) {
        set_error("fhdx_ctx_set_chemistry: nreaction must be 1..FHDX_MAX_REACTION and equal fhdx_params.nreaction");
        return FHDX_ERR_ARG;
    }
    if (ch->reaction_type != 0 && ch->reaction_type != 1) {  // chemistry_functions.cpp:131-133
        set_error("fhdx_ctx_set_chemistry: reaction_type must be 0 (deterministic) or 1 (CLE)");
        return FHDX_ERR_UNSUPPORTED;
    }
    if (!(ch->T0_chem > 0.)) {  // chemistry_functions.cpp:135
        set_error("fhdx_ctx_set_chemistry: T0_chem > 0 expected");
        return FHDX_ERR_ARG;
    }
    DevChem d{};
    d.nreaction = nr;
    d.reaction_type = ch->reaction_type;
    for (int m = 0; m < nr; ++m) {
        for (int n = 0; n < NS; ++n) {
            const int R = ch->stoich_R[m][n], Pp = ch->stoich_P[m][n];
            if (R < 0 || Pp < 0 || R > 100 || Pp > 100) {
                set_error("fhdx_ctx_set_chemistry: stoichiometric coefficients must be 0..100");
                return FHDX_ERR_ARG;
            }
            d.R[m][n] = (signed char)R;
            d.PR[m][n] = (signed char)(Pp - R);
        }
        d.rate[m] = ch->rate_const[m];
        d.alpha[m] = ch->alpha_param[m];
        d.beta[m] = ch->beta_param[m];
    }
    d.T0 = ch->T0_chem;
    d.Navo = c->p.Runiv / c->p.k_B;
    c->dp.chem = d;
    c->chem_set = true;
    return FHDX_OK;
}

int fhdx_chemistry_source(fhdx_ctx* c, fhdx_fab prim, fhdx_fab source, fhdx_fab ranchem, double dt, void* stream)
{
    if (!c || !(dt > 0.)) return FHDX_ERR_ARG;
    if (!c->chem_set) {
        set_error("fhdx_chemistry_source: call fhdx_ctx_set_chemistry first");
        return FHDX_ERR_ARG;
    }
    Box b;
    if (!cell_box(c, prim, c->dp.nprim, "prim", b) || !check_cell(c, b, source, c->dp.nvars, "source") ||
        !check_cell(c, b, ranchem, c->p.nreaction, "ranchem"))
        return FHDX_ERR_ARG;
    return ref_chem_source(box_params(c, b), make_arr(prim), make_arr(source), make_arr(ranchem), dt, S(stream));
}

int fhdx_fill_chem_noise(fhdx_ctx* c, uint64_t seed, uint64_t step, uint32_t instance, fhdx_fab ranchemA,
                         fhdx_fab ranchemB, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    if (!c->chem_set) {
        set_error("fhdx_fill_chem_noise: call fhdx_ctx_set_chemistry first");
        return FHDX_ERR_ARG;
    }
    Box b;
    if (!cell_box(c, ranchemA, c->p.nreaction, "ranchemA", b) ||
        !check_cell(c, b, ranchemB, c->p.nreaction, "ranchemB"))
        return FHDX_ERR_ARG;
    return ref_fill_chem_noise(box_params(c, b), make_noise_key(seed, step, instance), make_arr(ranchemA),
                               make_arr(ranchemB), S(stream));
}

// ---- evaluateStats ----------------------------------------------------------------------

int fhdx_evaluate_stats(fhdx_ctx* c, fhdx_fab cons, fhdx_fab consMean, fhdx_fab consVar, fhdx_fab prim,
                        fhdx_fab primMean, fhdx_fab primVar, fhdx_fab spatialCross, fhdx_fab miscStats,
                        double* miscVals, int steps, int cross_cell, void* stream)
{
    if (!c || !miscVals) return FHDX_ERR_ARG;
    const int nv = c->dp.nvars, np = c->dp.nprim;
    if (!check_cell(c, cons, nv, "cons") || !check_cell(c, consMean, nv, "consMean") ||
        !check_cell(c, consVar, nv, "consVar") || !check_cell(c, prim, np, "prim") ||
        !check_cell(c, primMean, np, "primMean") || !check_cell(c, primVar, np + 5, "primVar") ||
        !check_cell(c, spatialCross, 8, "spatialCross") || !check_cell(c, miscStats, 6, "miscStats"))
        return FHDX_ERR_ARG;
    if (steps < 1 || cross_cell < 0 || cross_cell >= c->p.n_cells[0]) {
        set_error("fhdx_evaluate_stats: steps must be >= 1 and 0 <= cross_cell < n_cells[0]");
        return FHDX_ERR_ARG;
    }
    const int n0 = c->p.n_cells[0];
    if (!c->st_yz) {
        const long nyz = (long)c->p.n_cells[1] * c->p.n_cells[2];
        c->st_nparts = (int)(nyz < 256 ? nyz : 256);
        FHDX_HIP_TRY(hipMalloc((void**)&c->st_yz, sizeof(double) * n0 * 19));
        FHDX_HIP_TRY(hipMalloc((void**)&c->st_part, sizeof(double) * c->st_nparts * n0 * 19));
        FHDX_HIP_TRY(hipMalloc((void**)&c->st_misc, sizeof(double) * 20));
    }
    StatsArgs a;
    a.cu = make_arr(cons);
    a.cm = make_arr(consMean);
    a.cv = make_arr(consVar);
    a.prim = make_arr(prim);
    a.pm = make_arr(primMean);
    a.pv = make_arr(primVar);
    a.sc = make_arr(spatialCross);
    a.ms = make_arr(miscStats);
    a.yz = c->st_yz;
    a.part = c->st_part;
    a.misc = c->st_misc;
    a.nparts = c->st_nparts;
    a.steps = steps;
    a.cross_cell = cross_cell;
    int r = stats_launch(c->dp, a, S(stream));
    if (r) return r;
    FHDX_HIP_TRY(hipMemcpyAsync(miscVals, c->st_misc, sizeof(double) * 20, hipMemcpyDeviceToHost, S(stream)));
    FHDX_HIP_TRY(hipStreamSynchronize(S(stream)));
    return FHDX_OK;
}

// ---- launch-shape tuning ---------------------------------------------------------------

namespace {

struct TuneKey {
    int ns, n0, n1, n2;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(ns, n0, n1, n2) < std::tie(o.ns, o.n0, o.n1, o.n2);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, fhdx_tune> g_tune_cache;

void apply_tune(FusedWorkspace& w, const fhdx_tune& t)
{
    w.tw_max = t.tile_width;
    w.rh = t.region_tiles;
    w.noise_inline = t.noise_inline != 0;
    w.kz = t.z_levels;
    w.face_bpc = t.face_blocks_per_cu;
    w.rec_in_update = t.records_in_update != 0;
}

bool valid_tune(const fhdx_tune& t)
{
    return (t.tile_width == 16 || t.tile_width == 32 || t.tile_width == 64) && t.region_tiles >= 1 &&
           t.region_tiles <= 64 && (t.noise_inline == 0 || t.noise_inline == 1) && t.z_levels >= 1 &&
           t.z_levels <= 64 && (t.face_blocks_per_cu == 0 || t.face_blocks_per_cu == 2 || t.face_blocks_per_cu == 3) &&
           (t.records_in_update == 0 || t.records_in_update == 1);
}

}  // namespace

int fhdx_ctx_set_tune(fhdx_ctx* c, const fhdx_tune* t)
{
    if (!c || !t || !valid_tune(*t)) {
        set_error("fhdx_ctx_set_tune: tile_width must be 16, 32 or 64, region_tiles and z_levels 1..64, "
                  "noise_inline 0/1, face_blocks_per_cu 0, 2 or 3, records_in_update 0/1");
        return FHDX_ERR_ARG;
    }
    apply_tune(c->fused, *t);
    return FHDX_OK;
}

int fhdx_ctx_get_tune(fhdx_ctx* c, fhdx_tune* t)
{
    if (!c || !t) return FHDX_ERR_ARG;
    t->tile_width = c->fused.tw_max;
    t->region_tiles = c->fused.rh;
    t->noise_inline = c->fused.noise_inline ? 1 : 0;
    t->z_levels = c->fused.kz;
    t->face_blocks_per_cu = c->fused.face_bpc;
    t->records_in_update = c->fused.rec_in_update ? 1 : 0;
    t->ms_per_step = 0.;
    return FHDX_OK;
}

int fhdx_ctx_autotune(fhdx_ctx* c, fhdx_fab cu, double dt, int trial_steps, fhdx_tune* best, fhdx_tune* trials,
                      int max_trials, int* n_trials, void* stream)
{
    if (!c || !best || trial_steps < 1 || !(dt > 0.)) return FHDX_ERR_ARG;
    if (n_trials) *n_trials = 0;
    if (!check_cell(c, cu, c->dp.nvars, "cu")) return FHDX_ERR_ARG;
    if (int r = reject_unfused(c, "fhdx_ctx_autotune")) return r;
    const TuneKey key{c->p.nspecies, c->p.n_cells[0], c->p.n_cells[1], c->p.n_cells[2]};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            apply_tune(c->fused, it->second);
            *best = it->second;
            return FHDX_OK;
        }
    }
    hipStream_t s = S(stream);
    // scratch copies: the caller's state is not advanced
    const size_t pts = (size_t)(cu.hi[0] - cu.lo[0] + 1) * (cu.hi[1] - cu.lo[1] + 1) * (cu.hi[2] - cu.lo[2] + 1);
    double *scu = nullptr, *sprim = nullptr;
    int* serr = nullptr;  // trial steps raise their flags here, not on the context
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = FHDX_OK;
    auto fail = [&](int code, const char* what) {
        if (!r) {
            r = code;
            set_error(std::string("fhdx_ctx_autotune: ") + what);
        }
    };
    if (hipMalloc((void**)&scu, pts * c->dp.nvars * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&sprim, pts * c->dp.nprim * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&serr, sizeof(int)) != hipSuccess)
        fail(FHDX_ERR_NOMEM, "scratch allocation");
    if (!r && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) fail(FHDX_ERR_HIP, "events");
    if (!r && hipMemcpyAsync(scu, cu.p, pts * c->dp.nvars * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess)
        fail(FHDX_ERR_HIP, "state copy");
    fhdx_fab fcu = cu, fprim = cu;
    fcu.p = scu;
    fprim.p = sprim;
    fprim.ncomp = c->dp.nprim;
    fhdx_tune top{};
    const bool timing_was_on = c->fused.timing.on;
    c->fused.timing.on = false;
    uint64_t step = 0;
    auto trial = [&](fhdx_tune& t) -> int {
        apply_tune(c->fused, t);
        float ms = 0.f;
        int q = fused_rk3_step(c->p, c->dp, c->fused, fcu, fprim, dt, make_noise_key(0, step++, 0), nullptr, nullptr,
                               nullptr, nullptr, serr, s);
        if (!q && hipEventRecord(e0, s) != hipSuccess) q = FHDX_ERR_HIP;
        for (int k = 0; !q && k < trial_steps; ++k)
            q = fused_rk3_step(c->p, c->dp, c->fused, fcu, fprim, dt, make_noise_key(0, step++, 0), nullptr, nullptr,
                               nullptr, nullptr, serr, s);
        if (!q && hipEventRecord(e1, s) != hipSuccess) q = FHDX_ERR_HIP;
        if (!q && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            q = FHDX_ERR_HIP;
        t.ms_per_step = ms / trial_steps;
        if (!q && trials && n_trials && *n_trials < max_trials) trials[(*n_trials)++] = t;
        return q;
    };
    // coordinate descent from the default shape: each parameter swept with the others fixed
    fhdx_tune cur{64, 1, 1, 1, 0, 0, 0.};
    if (!r) {
        int q = trial(cur);
        if (q) fail(q, "trial step failed");
    }
    top = cur;
    const int nz = c->p.n_cells[2];
    for (int dim = 0; dim < 6 && !r; ++dim) {
        std::vector<int> vals;
        if (dim == 0) vals = {64, 32, 16};
        if (dim == 1) vals = {1, 2, 4, 8};
        if (dim == 2) vals = {1, 2, 4, 8, 16};
        if (dim == 3) vals = {1, 0};
        if (dim == 4) vals = {0, 3, 2};
        if (dim == 5) vals = {0, 1};
        for (int v : vals) {
            fhdx_tune t = top;
            int32_t* field = dim == 0 ? &t.tile_width : dim == 1 ? &t.region_tiles : dim == 2 ? &t.z_levels
                           : dim == 3 ? &t.noise_inline : dim == 4 ? &t.face_blocks_per_cu
                                                                  : &t.records_in_update;
            if (*field == v) continue;
            if (dim == 0 && v > 16 && (v >> 1) >= c->p.n_cells[0]) continue;  // same tiles as a narrower one
            if (dim == 2 && v > nz) continue;
            *field = v;
            int q = trial(t);
            if (q) {
                fail(q, "trial step failed");
                break;
            }
            if (t.ms_per_step < top.ms_per_step) top = t;
        }
    }
    c->fused.timing.on = timing_was_on;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    if (scu) (void)hipFree(scu);
    if (sprim) (void)hipFree(sprim);
    if (serr) (void)hipFree(serr);
    if (r) return r;
    apply_tune(c->fused, top);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = top;
    }
    *best = top;
    return FHDX_OK;
}

int fhdx_ctx_timing_enable(fhdx_ctx* c, int enable)
{
    if (!c) return FHDX_ERR_ARG;
    c->fused.timing.reset();
    c->fused.timing.on = enable != 0;
    return FHDX_OK;
}

int fhdx_timing_kinds(void) { return KT_COUNT; }

int fhdx_ctx_timing_read(fhdx_ctx* c, int kind, char* name, int name_len, int64_t* launches, double* total_ms)
{
    if (!c || kind < 0 || kind >= KT_COUNT) return FHDX_ERR_ARG;
    int r = c->fused.timing.resolve();
    if (r) {
        set_error("waiting for timing events failed");
        return r;
    }
    if (name && name_len > 0) {
        std::strncpy(name, kKernelKindNames[kind], (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (launches) *launches = c->fused.timing.launches[kind];
    if (total_ms) *total_ms = c->fused.timing.total_ms[kind];
    return FHDX_OK;
}

}  // extern "C"
