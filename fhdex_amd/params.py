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
        ("t_lo", C.c_double * 3),
        ("t_hi", C.c_double * 3),
        ("bc_Yk", (C.c_double * MAX_SPECIES) * 6),
        ("bc_Xk", (C.c_double * MAX_SPECIES) * 6),
        ("p_lo", C.c_double * 3),
        ("p_hi", C.c_double * 3),
        ("rho_lo", C.c_double * 3),
        ("rho_hi", C.c_double * 3),
    ]

    @property
    def periodic(self):
        return tuple(self.bc_vel_lo[d] == -1 for d in range(3))

    @property
    def has_walls(self) -> bool:
        return not all(self.periodic)

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
    """A run: parameters plus the initial state (rho0, mass fractions rhobar, T_init and the
    InitConsVar prob_type) and time step."""

    def __init__(self, name, params: Params, rho0, rhobar, T_init, dt, nsteps, chem=None, prob_type=1):
        self.name = name
        self.chem = chem
        self.prob_type = prob_type  # InitConsVar's initial state (compressible_functions.cpp:103-413)
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


def deterministic_wave(nsteps=100000):
    """exec/compressible/inputs_test_deterministic: two identical argon species (M 39.928...,
    d 3.66e-8 cm), T 300 K, rho0 1.78e-3, a 4 x 80 x 4 periodic box of 3.125e-6 cm cells, dt
    1e-12 s, deterministic fluxes (stoch_stress_form 0), prob_type 100: density
    rho0 (1 + 0.1 sin(2 pi y / L_y)) at the uniform energy density, an isobaric (entropy) wave."""
    M = 39.92820856716215701081
    p = make_params(2, (4, 80, 4), (1.25e-5 / 4, 2.5e-4 / 80, 1.25e-5 / 4), [M, M], [3.66e-8, 3.66e-8],
                    stoch_stress_form=0)
    return Config("deterministic_wave_4x80x4", p, 1.78e-3, [0.5, 0.5], 300.0, 1.0e-12, nsteps, prob_type=100)


def taylor_green(n=64, nsteps=100):
    """The Taylor-Green vortex deck (exec/compressible/inputs_tg, its final block): two
    argon species (M 39.948, d 3.63e-8 cm, Y 1/2 each), T 273 K, an n^3 periodic box of edge
    7.85e-3 cm (64^3 cells), dt 1e-9 s, 100 steps, prob_type 4 (velocity amplitude 9237 cm/s,
    mean pressure 884147 dyn/cm^2; the deck's turbForcing only adds energy-spectrum
    diagnostics in main_driver.cpp)."""
    p = make_params(2, (n, n, n), 0.00785 / 64, [39.948, 39.948], [3.63e-8, 3.63e-8])
    rho = 39.948 / (R_UNIV / K_B) * 884.147e3 / (K_B * 273.0)
    return Config(f"taylor_green_{n}^3", p, rho, [0.5, 0.5], 273.0, 1.0e-9, nsteps, prob_type=4)


# Eight hard-sphere species: the four of exec/compressible/inputs_profiling_gpu:82-83 (He, Ne,
# Ar, Kr), then Xe, N2, O2 and CO2 with rigid-sphere diameters of the size room-temperature
# viscosities give (approximate values: they only have to be distinct and physical).
EIGHT_SPECIES_M = [4.0026, 20.1797, 39.9480, 83.8000, 131.293, 28.0134, 31.9988, 44.0095]
EIGHT_SPECIES_D = [2.18e-8, 2.58e-8, 3.63e-8, 4.16e-8, 4.85e-8, 3.75e-8, 3.61e-8, 4.59e-8]


def species_prefix(n_species, edge, nsteps, shape=None):
    """Species 1..n_species of the eight-species list with equal mass fractions (1..4 is the
    served request mix); n_species == 1 is written as two identical species (argon form).
    The box is edge^3 cells, or `shape`."""
    if n_species <= 1:
        return argon_gas(edge, nsteps, shape)
    if n_species > len(EIGHT_SPECIES_M):
        raise ValueError(f"species_prefix: at most {len(EIGHT_SPECIES_M)} species (MAX_SPECIES)")
    M = EIGHT_SPECIES_M[:n_species]
    d = EIGHT_SPECIES_D[:n_species]
    box = f"{edge}^3" if shape is None else "x".join(str(s) for s in shape)
    shape = tuple(shape) if shape is not None else (edge, edge, edge)
    p = make_params(n_species, shape, 8e-6, M, d)
    rhobar = [1.0 / n_species] * n_species
    return Config(f"species{n_species}_{box}", p, 4.83e-4, rhobar, 300.0, MIXTURE_DT, nsteps)


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


def set_walls(p: Params, bc_vel, bc_therm, bc_mass, t_lo=(0.0, 0.0, 0.0), t_hi=(0.0, 0.0, 0.0),
              bc_Yk=None, p_lo=(0.0, 0.0, 0.0), p_hi=(0.0, 0.0, 0.0), rho_lo=(0.0, 0.0, 0.0),
              rho_hi=(0.0, 0.0, 0.0)) -> Params:
    """Boundary codes of the reference's namelist (common_namespace.H): bc_vel, bc_therm and
    bc_mass are ((lo, hi) per direction) or a single (lo, hi) pair or code for all three.
    A direction with bc_vel -1 is periodic and every code of it becomes -1 (SetupBC,
    boundary.cpp:3-12).  bc_Yk is one composition for every bc_mass 2 face or a list of six
    (face 2*direction + side); the mole fractions follow from GetMolfrac as SetupCWall does
    (boundary.cpp:14-170, compressible_functions.H:139-152)."""

    def per_dir(v):
        if isinstance(v, int):
            return [(v, v)] * 3
        if len(v) == 2 and isinstance(v[0], int):
            return [tuple(v)] * 3
        return [tuple(x) if not isinstance(x, int) else (x, x) for x in v]

    vel, therm, mass = per_dir(bc_vel), per_dir(bc_therm), per_dir(bc_mass)
    for d in range(3):
        if vel[d][0] == -1:
            vel[d] = therm[d] = mass[d] = (-1, -1)
        p.bc_vel_lo[d], p.bc_vel_hi[d] = vel[d]
        p.bc_therm_lo[d], p.bc_therm_hi[d] = therm[d]
        p.bc_mass_lo[d], p.bc_mass_hi[d] = mass[d]
        p.t_lo[d] = float(t_lo[d])
        p.t_hi[d] = float(t_hi[d])
        p.p_lo[d], p.p_hi[d] = float(p_lo[d]), float(p_hi[d])
        p.rho_lo[d], p.rho_hi[d] = float(rho_lo[d]), float(rho_hi[d])
    ns = p.nspecies
    for f in range(6):
        code = mass[f // 2][f % 2]
        for n in range(MAX_SPECIES):
            p.bc_Yk[f][n] = 0.0
            p.bc_Xk[f][n] = 0.0
        if code != 2 and code != 3:
            continue
        Y = bc_Yk if (bc_Yk is not None and not hasattr(bc_Yk[0], "__len__")) else bc_Yk[f]
        for n in range(ns):
            p.bc_Yk[f][n] = float(Y[n])
        molmix = 0.0
        for n in range(ns):
            molmix += p.bc_Yk[f][n] / p.molmass[n]
        molmix = 1.0 / molmix
        for n in range(ns):
            p.bc_Xk[f][n] = p.bc_Yk[f][n] * molmix / p.molmass[n]
        if code == 3:  # reservoir density from p, T and the composition (GetDensity), if not given
            tw = p.t_hi[f // 2] if f % 2 else p.t_lo[f // 2]
            pw = p.p_hi[f // 2] if f % 2 else p.p_lo[f // 2]
            if (p.rho_hi[f // 2] if f % 2 else p.rho_lo[f // 2]) <= 0.0 and pw > 0.0 and tw > 0.0:
                rho = pw / (p.Runiv / molmix) / tw
                if f % 2:
                    p.rho_hi[f // 2] = rho
                else:
                    p.rho_lo[f // 2] = rho
    return p


def ambient_pressure(p: Params, rho0, rhobar, T) -> float:
    """GetPressureGas of the initial state (compressible_functions.H:186-199): the reference's
    default reservoir pressure (boundary.cpp:43-47)."""
    molmix = 0.0
    for n in range(p.nspecies):
        molmix += rhobar[n] / p.molmass[n]
    molmix = 1.0 / molmix
    return rho0 * (p.Runiv / molmix) * T


def wall_deck(name, n=None, nsteps=3):
    """The reference's wall decks (exec/compressible/inputs_regression_*, inputs_giantfluct_3d).

    The first three hold four identical argon species (M 39.948, d 3.66e-8 cm, Y 1/4 each),
    rho 1.78e-3, T_init 400 K, and start uniform (prob_type 1):

    * slip_wall_adiabatic: 32^3 cells of 6.25e-5/32 cm, dt 2e-12 s, every face a slip,
      adiabatic, species-wall face (codes 1).
    * noslip_reservoir_isothermal: 32^3 cells of 1e-4/32 cm, dt 2e-12 s, no-slip isothermal
      walls at 390 K (lo) / 410 K (hi) with a fixed composition Y = 0.22/0.24/0.26/0.28 (codes 2).
    * temp_grad: 80 x 10 x 10 cells of 2e-6 cm, dt 1e-12 s, advection_type 1; x: no-slip
      isothermal walls at 280.8 / 508.5 K, y and z: slip adiabatic walls; species walls.

    * RT_3d (inputs_regression_RT_3d): 64 x 32 x 32 cells of 1.2e-5 cm, dt 5e-12 s, four
      species (M 2/20/2/20, d 2/20/20/2e-8 cm), rho0 4.83e-4, T 300 K, gravity -4e12 cm/s^2 in z,
      slip adiabatic species walls in z; prob_type 2 (two layers at hydrostatic equilibrium).
    * giantfluct_3d (inputs_giantfluct_3d): 256 x 128 x 1 cells of 2.7e-5 cm, dt 1e-10 s, three
      argon-mass species (d 3.63/3.63/36.3e-8 cm), rho0 4.83e-4, T 300 K, no-slip isothermal
      (300 K) walls of fixed composition in y; prob_type 3 (mass fractions linear between the
      two wall compositions).  The box is one cell deep in z (quasi-2-D), which the
      reference-shaped step serves; the fused step needs at least 4 cells per direction.

    n rescales the box (cells per direction; the x count of temp_grad; the y count of RT_3d
    and giantfluct_3d, whose x count is 2 n) at the deck's cell size, for runs small enough
    for the oracle."""
    M = [39.9480] * 4
    dmol = [3.66e-8] * 4
    if name == "slip_wall_adiabatic":
        cells = (n or 32,) * 3
        p = make_params(4, cells, 6.25e-5 / 32, M, dmol)
        set_walls(p, 1, 1, 1, t_lo=(300.0,) * 3, t_hi=(300.0,) * 3)
        dt = 2e-12
    elif name == "noslip_reservoir_isothermal":
        cells = (n or 32,) * 3
        p = make_params(4, cells, 1e-4 / 32, M, dmol)
        set_walls(p, 2, 2, 2, t_lo=(390.0,) * 3, t_hi=(410.0,) * 3, bc_Yk=[0.22, 0.24, 0.26, 0.28])
        dt = 2e-12
    elif name == "temp_grad":
        cells = (n or 80, 10, 10)
        p = make_params(4, cells, 2e-6, M, dmol, advection_type=1)
        set_walls(p, [(2, 2), (1, 1), (1, 1)], [(2, 2), (1, 1), (1, 1)], 1,
                  t_lo=(280.8, 300.0, 300.0), t_hi=(508.5, 400.0, 400.0))
        dt = 1e-12
    elif name == "RT_3d":
        m = n or 32
        cells = (2 * m, m, m)
        p = make_params(4, cells, 7.68e-4 / 64, [2.0, 20.0, 2.0, 20.0], [2.0e-8, 20.0e-8, 20.0e-8, 2.0e-8],
                        grav=(0.0, 0.0, -4.0e12))
        set_walls(p, [(-1, -1), (-1, -1), (1, 1)], 1, 1, t_lo=(300.0,) * 3, t_hi=(300.0,) * 3)
        return Config(f"{name}_{'x'.join(map(str, cells))}", p, 4.83e-4, [0.25] * 4, 300.0, 5.0e-12, nsteps,
                      prob_type=2)
    elif name == "giantfluct_3d":
        m = n or 128
        cells = (2 * m, m, 1)
        p = make_params(3, cells, (6.912e-3 / 256, 3.456e-3 / 128, 2.7e-5), [39.9480] * 3,
                        [3.63e-8, 3.63e-8, 36.3e-8])
        set_walls(p, [(-1, -1), (2, 2), (-1, -1)], 2, 2, t_lo=(300.0,) * 3, t_hi=(300.0,) * 3,
                  bc_Yk=[[0.0] * 3, [0.0] * 3, [0.2, 0.09316672, 0.70683328],
                         [0.3, 0.40683328, 0.29316672], [0.0] * 3, [0.0] * 3])
        return Config(f"{name}_{'x'.join(map(str, cells))}", p, 4.83e-4, [0.25, 0.25, 0.5], 300.0, 1.0e-10,
                      nsteps, prob_type=3)
    else:
        raise ValueError(name)
    return Config(f"{name}_{'x'.join(map(str, cells))}", p, 1.78e-3, [0.25] * 4, 400.0, dt, nsteps)
