"""ctypes mirror of ``include/fhdx.h`` plus the run configurations of the compressible path.

The structs here are the C ABI's data formats (``fhdx_params``, ``fhdx_fab``); the
configuration builders restate the reference's input decks
(``exec/compressible/inputs_profiling_gpu``, ``inputs_test_deterministic``) and the defaults
of ``src_common/common_functions.cpp:320-490``.
"""
from __future__ import annotations

import ctypes as C
import math

MAX_SPECIES = 8
MAX_REACTION = 8

# physical constants, src_common/common_functions.cpp:431,433
K_B = 1.38064852e-16
R_UNIV = 8.314462175e7


class Params(C.Structure):
    """``fhdx_params`` (include/fhdx.h)."""

    _fields_ = [
        ("nspecies", C.c_int32),
        ("n_cells", C.c_int32 * 3),
        ("ngc", C.c_int32),
        ("algorithm_type", C.c_int32),
        ("advection_type", C.c_int32),
        ("visc_type", C.c_int32),
        ("transport_type", C.c_int32),
        ("stoch_stress_form", C.c_int32),
        ("nreaction", C.c_int32),
        ("bc_vel_lo", C.c_int32 * 3),
        ("bc_vel_hi", C.c_int32 * 3),
        ("bc_mass_lo", C.c_int32 * 3),
        ("bc_mass_hi", C.c_int32 * 3),
        ("bc_therm_lo", C.c_int32 * 3),
        ("bc_therm_hi", C.c_int32 * 3),
        ("pad0", C.c_int32),
        ("dx", C.c_double * 3),
        ("molmass", C.c_double * MAX_SPECIES),
        ("diameter", C.c_double * MAX_SPECIES),
        ("dof", C.c_double * MAX_SPECIES),
        ("hcv", C.c_double * MAX_SPECIES),
        ("hcp", C.c_double * MAX_SPECIES),
        ("e0", C.c_double * MAX_SPECIES),
        ("k_B", C.c_double),
        ("Runiv", C.c_double),
        ("variance_coef_mom", C.c_double),
        ("variance_coef_ener", C.c_double),
        ("variance_coef_mass", C.c_double),
        ("grav", C.c_double * 3),
        ("zeta_ratio", C.c_double),
    ]

    @property
    def nvars(self) -> int:
        return 5 + self.nspecies

    @property
    def nprimvars(self) -> int:
        return 6 + 2 * self.nspecies

    @property
    def shape(self):
        return tuple(self.n_cells)

    def copy(self) -> "Params":
        q = Params()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(Params))
        return q


class Chemistry(C.Structure):
    """``fhdx_chemistry`` (include/fhdx.h): reactions of chemistry_namespace.H:1-39."""

    _fields_ = [
        ("nreaction", C.c_int32),
        ("reaction_type", C.c_int32),
        ("stoich_R", (C.c_int32 * MAX_SPECIES) * MAX_REACTION),
        ("stoich_P", (C.c_int32 * MAX_SPECIES) * MAX_REACTION),
        ("rate_const", C.c_double * MAX_REACTION),
        ("alpha_param", C.c_double * MAX_REACTION),
        ("beta_param", C.c_double * MAX_REACTION),
        ("T0_chem", C.c_double),
    ]


def make_chemistry(stoich_R, stoich_P, rate_const, T0_chem, reaction_type=1, alpha=None,
                   beta=None) -> Chemistry:
    """Reaction data as InitializeChemistryNamespace reads it (chemistry_functions.cpp:42-127):
    stoich_R[m][n], stoich_P[m][n] per reaction m and species n."""
    c = Chemistry()
    c.nreaction = len(rate_const)
    c.reaction_type = reaction_type
    for m in range(c.nreaction):
        for n in range(len(stoich_R[m])):
            c.stoich_R[m][n] = int(stoich_R[m][n])
            c.stoich_P[m][n] = int(stoich_P[m][n])
        c.rate_const[m] = float(rate_const[m])
        c.alpha_param[m] = float(alpha[m]) if alpha is not None else 0.0
        c.beta_param[m] = float(beta[m]) if beta is not None else 0.0
    c.T0_chem = float(T0_chem)
    return c


class Fab(C.Structure):
    """``fhdx_fab``: data pointer and the inclusive index box it covers."""

    _fields_ = [
        ("p", C.c_void_p),
        ("lo", C.c_int32 * 3),
        ("hi", C.c_int32 * 3),
        ("ncomp", C.c_int32),
        ("pad", C.c_int32),
    ]


def make_params(nspecies: int, n_cells, dx, molmass, diameter, *, ngc: int = 2,
                dof=None, e0=None, zeta_ratio: float = -1.0, visc_type: int = 2,
                advection_type: int = 2, algorithm_type: int = 2,
                stoch_stress_form: int = 1, variance_coef=(1.0, 1.0, 1.0),
                grav=(0.0, 0.0, 0.0)) -> Params:
    """Periodic run parameters with hcv/hcp from dof as GetHcGas does
    (src_compressible/compressible_functions.cpp:91-100)."""
    p = Params()
    p.nspecies = nspecies
    for d in range(3):
        p.n_cells[d] = int(n_cells[d])
        p.dx[d] = float(dx[d] if hasattr(dx, "__len__") else dx)
        p.bc_vel_lo[d] = p.bc_vel_hi[d] = -1
        p.bc_mass_lo[d] = p.bc_mass_hi[d] = -1
        p.bc_therm_lo[d] = p.bc_therm_hi[d] = -1
        p.grav[d] = float(grav[d])
    p.ngc = ngc
    p.algorithm_type = algorithm_type
    p.advection_type = advection_type
    p.visc_type = visc_type
    p.transport_type = 1
    p.stoch_stress_form = stoch_stress_form
    p.nreaction = 0
    p.k_B = K_B
    p.Runiv = R_UNIV
    p.variance_coef_mom, p.variance_coef_ener, p.variance_coef_mass = variance_coef
    p.zeta_ratio = zeta_ratio
    for n in range(MAX_SPECIES):
        p.molmass[n] = 1.0
        p.diameter[n] = 1.0
        p.dof[n] = 3.0
        p.e0[n] = 0.0
    for n in range(nspecies):
        p.molmass[n] = float(molmass[n])
        p.diameter[n] = float(diameter[n])
        if dof is not None:
            p.dof[n] = float(dof[n])
        if e0 is not None:
            p.e0[n] = float(e0[n])
        p.hcv[n] = 0.5 * p.dof[n] * R_UNIV / p.molmass[n]
        p.hcp[n] = 0.5 * (2.0 + p.dof[n]) * R_UNIV / p.molmass[n]
    return p


class Config:
    """A run: parameters plus the uniform initial state and time step."""

    def __init__(self, name, params: Params, rho0, rhobar, T_init, dt, nsteps, chem=None):
        self.name = name
        self.chem = chem
        self.params = params
        self.rho0 = rho0
        self.rhobar = list(rhobar)
        self.T_init = T_init
        self.dt = dt
        self.nsteps = nsteps

    @property
    def cells(self) -> int:
        n = self.params.n_cells
        return n[0] * n[1] * n[2]


def sound_speed(params: Params, rhobar, T) -> float:
    """GetSoundSpeed (compressible_functions.H:201-219) for mass fractions rhobar."""
    molmix = sum(rhobar[n] / params.molmass[n] for n in range(params.nspecies))
    molmix = 1.0 / molmix
    cv = sum(rhobar[n] * params.hcv[n] for n in range(params.nspecies))
    cp = sum(rhobar[n] * params.hcp[n] for n in range(params.nspecies))
    return math.sqrt(cp / cv * (R_UNIV / molmix) * T)


def cfl_dt(params: Params, rhobar, T, cfl=0.1) -> float:
    """dt = cfl * dx / c_s at the uniform start (the reference uses a fixed dt,
    main_driver.cpp:112)."""
    return cfl * min(params.dx) / sound_speed(params, rhobar, T)


# Acoustic CFL of the argon runs.  At CFL 0.2 (dt = 5.2e-11 s) the explicit heat conduction
# puts the 3-D checkerboard mode at 12 kappa dt/(rho c_v dx^2) = 2.85 > 2.51, outside RK3's
# stability interval, and noise-driven runs blow up within ~15 steps; CFL 0.1 (dt = 2.6e-11 s)
# keeps a factor ~1.8 of margin.
ARGON_CFL = 0.1


def argon_gas(n, nsteps=100, shape=None):
    """Hard-sphere argon written as two identical species (the reference's single-gas form,
    exec/compressible/inputs_test_deterministic:81-87): M 39.95, d 3.66e-8 cm, T 273 K,
    rho 1.78e-3 g/cm^3, dx 8e-6 cm, acoustic CFL 0.1 (see ARGON_CFL).  The box is n^3 cells,
    or `shape` = (nx, ny, nz)."""
    shape = tuple(shape) if shape is not None else (n, n, n)
    p = make_params(2, shape, 8e-6, [39.95, 39.95], [3.66e-8, 3.66e-8])
    rhobar = [0.5, 0.5]
    dt = cfl_dt(p, rhobar, 273.0, ARGON_CFL)
    return Config(f"argon2_{shape[0]}x{shape[1]}x{shape[2]}", p, 1.78e-3, rhobar, 273.0, dt, nsteps)


# The reference's own four-species deck runs at a fixed dt of 1e-12 s
# (exec/compressible/inputs_profiling_gpu:18).  An acoustic CFL of 0.2 (dt = 2.7e-11 s) is
# not usable there: the explicit thermal-diffusion number kappa dt/(rho c_p dx^2) would be
# ~0.43, above the ~0.21 that RK3 tolerates in 3-D, and the state blows up in one step.
MIXTURE_DT = 1e-12


def four_species(n, nsteps=200):
    """Four hard-sphere species of exec/compressible/inputs_profiling_gpu:51-52,82-83:
    M 4.0026/20.1797/39.948/83.8, d 2.18/2.58/3.63/4.16e-8, Y 1/4 each, T 300 K,
    rho 4.83e-4, dx 8e-6 cm, dt 1e-12 s (the deck's fixed_dt)."""
    M = [4.0026, 20.1797, 39.9480, 83.8000]
    d = [2.18e-8, 2.58e-8, 3.63e-8, 4.16e-8]
    p = make_params(4, (n, n, n), 8e-6, M, d)
    rhobar = [0.25] * 4
    return Config(f"species4_{n}^3", p, 4.83e-4, rhobar, 300.0, MIXTURE_DT, nsteps)


def species_prefix(n_species, edge, nsteps):
    """Species 1..n_species of the four-species list with equal mass fractions (the served
    request mix); n_species == 1 is written as two identical species (argon form)."""
    if n_species <= 1:
        return argon_gas(edge, nsteps)
    M = [4.0026, 20.1797, 39.9480, 83.8000][:n_species]
    d = [2.18e-8, 2.58e-8, 3.63e-8, 4.16e-8][:n_species]
    p = make_params(n_species, (edge, edge, edge), 8e-6, M, d)
    rhobar = [1.0 / n_species] * n_species
    return Config(f"species{n_species}_{edge}^3", p, 4.83e-4, rhobar, 300.0, MIXTURE_DT, nsteps)


def no2_n2o4(n, nsteps=100, reaction_type=1):
    """2 NO2 <-> N2O4 with the chemical Langevin source, from the reference's deck
    exec/compressible_stag/test_NO2_dimerization_eq/inputs_NO2_dimerization_eq (the same
    compute_compressible_chemistry_source_CLE the collocated scheme calls): M 46.0055 / 92.0110,
    d 3.765e-8 / 4.621e-8 cm, dof 7.279586 / 18.049072, e0 4.685624e9 / -1.730782e9, Y 0.723969
    / 0.276031, T 350 K, rho 1.858338e-3, k_B 1.38065e-16, Runiv 8.31446e7, dx 8e-6 cm, dt 1e-12
    s; rate constants 4.065753e11 / 6.236899e7, alpha -5.337990e10 / 5.37e11, beta 0.644950 /
    -1.1, T0_chem 350.  (The collocated deck exec/compressible/chemistry_inputs_files/
    inputs_NO2_N2O4_eq_SF gives rate constants in other units: its source is ~1e24 g/cm^3/s at
    its own composition.)"""
    p = make_params(2, (n, n, n), 1.28e-4 / 16, [46.0055, 92.0110], [3.765e-08, 4.621e-08],
                    dof=[7.279586, 18.049072], e0=[4.685624e+09, -1.730782e+09])
    p.k_B = 1.38065e-16
    p.Runiv = 8.31446e7
    for q in range(2):
        p.hcv[q] = 0.5 * p.dof[q] * p.Runiv / p.molmass[q]
        p.hcp[q] = 0.5 * (2.0 + p.dof[q]) * p.Runiv / p.molmass[q]
    p.nreaction = 2
    chem = make_chemistry([[2, 0], [0, 1]], [[0, 1], [2, 0]], [4.065753e+11, 6.236899e+07],
                          350.0, reaction_type=reaction_type, alpha=[-5.337990e+10, 5.370000e+11],
                          beta=[0.644950, -1.100000])
    return Config(f"no2_n2o4_{n}^3", p, 1.858338e-03, [0.723969, 0.276031], 350.0, 1e-12,
                  nsteps, chem=chem)
