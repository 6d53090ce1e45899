"""The C ABI without a GPU: the library loads, exports every symbol include/fhdx.h declares,
the struct layouts agree with the header, and the pure host functions behave."""
import ctypes as C
import os
import re

import pytest

from fhdex_amd import lib as L
from fhdex_amd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "fhdx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fhdx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = header_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert sorted(L.EXPORTS) == sorted(names)


def test_abi_version_and_status_strings():
    lib = L.lib()
    assert lib.fhdx_abi_version() == 2
    assert lib.fhdx_status_string(0) == b"ok"
    assert b"mass fraction" in lib.fhdx_status_string(4)


def test_struct_layouts_match_header(tmp_path):
    """ctypes offsets equal the C compiler's offsetof for every field of both structs."""
    fields = [f for f, _ in P.Params._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "fhdx.h"', "int main(void){"]
    src.append('printf("%zu %zu\\n", sizeof(fhdx_params), sizeof(fhdx_fab));')
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof(fhdx_params, {f}));')
    for f, _ in P.Fab._fields_:
        src.append(f'printf("%zu\\n", offsetof(fhdx_fab, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split()
    vals = [int(v) for v in out]
    assert vals[0] == C.sizeof(P.Params) and vals[1] == C.sizeof(P.Fab)
    offs = vals[2:]
    for i, f in enumerate(fields):
        assert getattr(P.Params, f).offset == offs[i], f
    for i, (f, _) in enumerate(P.Fab._fields_):
        assert getattr(P.Fab, f).offset == offs[len(fields) + i], f


def test_serving_and_tune_struct_layouts_match_header(tmp_path):
    """fhdx_request / fhdx_completion / fhdx_tune: ctypes offsets equal offsetof."""
    from fhdex_amd import serving as S
    structs = [("fhdx_request", S.Request), ("fhdx_completion", S.Completion), ("fhdx_tune", L.Tune)]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "fhdx.h"', "int main(void){"]
    for cname, py in structs:
        src.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout2.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout2"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)])
    vals = [int(v) for v in subprocess.check_output([str(exe)]).decode().split()]
    q = 0
    for cname, py in structs:
        assert vals[q] == C.sizeof(py), cname
        q += 1
        for f, _ in py._fields_:
            assert getattr(py, f).offset == vals[q], (cname, f)
            q += 1


def test_params_default_finalize_check():
    lib = L.lib()
    p = P.Params()
    lib.fhdx_params_default(C.byref(p))
    assert p.ngc == 2 and p.advection_type == 2 and p.algorithm_type == 2
    assert p.hcv[0] < 0
    assert lib.fhdx_params_check(C.byref(p)) == 1          # hcv not derived yet
    p.molmass[0] = p.molmass[1] = 39.95
    p.diameter[0] = p.diameter[1] = 3.66e-8
    lib.fhdx_params_finalize(C.byref(p))
    assert p.hcv[0] == pytest.approx(1.5 * P.R_UNIV / 39.95)
    assert lib.fhdx_params_check(C.byref(p)) == 0
    t = p.copy()
    t.transport_type = 3                                   # Hirschfelder-Curtiss-Bird: binary only
    assert lib.fhdx_params_check(C.byref(t)) == 0
    t.nspecies = 3
    t.molmass[2], t.diameter[2] = 39.95, 3.66e-8
    lib.fhdx_params_finalize(C.byref(t))
    assert lib.fhdx_params_check(C.byref(t)) == 2
    assert b"binary" in lib.fhdx_last_error()
    t = p.copy()
    t.transport_type = 2                                   # Valk-Waldmann: any species count
    assert lib.fhdx_params_check(C.byref(t)) == 0
    t.transport_type = 4
    assert lib.fhdx_params_check(C.byref(t)) == 2
    t.transport_type = 0
    assert lib.fhdx_params_check(C.byref(t)) == 2
    q = p.copy()
    q.bc_vel_lo[0] = 1
    assert lib.fhdx_params_check(C.byref(q)) == 1          # one side periodic, the other not
    assert b"periodic" in lib.fhdx_last_error()
    q = p.copy()
    P.set_walls(q, [(1, 1), (-1, -1), (-1, -1)], 1, 1)
    assert lib.fhdx_params_check(C.byref(q)) == 0          # slip adiabatic species walls in x
    r = q.copy()
    r.bc_mass_hi[0] = 3                                    # x reservoir without its gas state
    assert lib.fhdx_params_check(C.byref(r)) == 1
    assert b"reservoir" in lib.fhdx_last_error()
    r = q.copy()                                           # x reservoir: rho follows from p, T, Y
    P.set_walls(r, [(2, 2), (-1, -1), (-1, -1)], [(2, 2), (-1, -1), (-1, -1)], [(3, 1), (-1, -1), (-1, -1)],
                t_lo=(300.0, 0.0, 0.0), t_hi=(300.0, 0.0, 0.0), bc_Yk=[0.5, 0.5], p_lo=(1.0e6, 0.0, 0.0))
    r.rho_lo[0] = 0.0
    lib.fhdx_params_finalize(C.byref(r))
    assert r.rho_lo[0] == pytest.approx(1.0e6 * 39.95 / (P.R_UNIV * 300.0))
    assert lib.fhdx_params_check(C.byref(r)) == 0
    r = q.copy()                                           # y reservoirs are not
    P.set_walls(r, [(-1, -1), (2, 2), (-1, -1)], [(-1, -1), (2, 2), (-1, -1)], [(-1, -1), (3, 3), (-1, -1)],
                t_lo=(0.0, 300.0, 0.0), t_hi=(0.0, 300.0, 0.0), bc_Yk=[0.5, 0.5], p_lo=(0.0, 1e6, 0.0),
                p_hi=(0.0, 1e6, 0.0))
    assert lib.fhdx_params_check(C.byref(r)) == 2
    assert b"x faces only" in lib.fhdx_last_error()
    r = q.copy()
    r.bc_therm_lo[0] = 2                                   # isothermal without a wall temperature
    assert lib.fhdx_params_check(C.byref(r)) == 1
    r = q.copy()
    P.set_walls(r, 2, 2, 2, t_lo=(390.0,) * 3, t_hi=(410.0,) * 3, bc_Yk=[0.6, 0.4])
    assert lib.fhdx_params_check(C.byref(r)) == 0
    r.bc_Xk[3][1] = 0.0                                    # the finalize step fills X from Y
    lib.fhdx_params_finalize(C.byref(r))
    assert r.bc_Xk[3][1] == pytest.approx(0.4) and lib.fhdx_params_check(C.byref(r)) == 0
    r.algorithm_type = 1
    assert lib.fhdx_params_check(C.byref(r)) == 2          # walls need the species ghosts
    q = p.copy()
    q.n_cells[2] = 1                                       # quasi-2-D: one periodic cell deep
    assert lib.fhdx_params_check(C.byref(q)) == 0
    q.n_cells[2] = 0
    assert lib.fhdx_params_check(C.byref(q)) == 1
    q = p.copy()
    P.set_walls(q, [(-1, -1), (-1, -1), (1, 1)], 1, 1)
    q.n_cells[2] = 3                                       # setBC mirrors ngc = 2 cells per wall
    assert lib.fhdx_params_check(C.byref(q)) == 2
    assert b"2 ngc" in lib.fhdx_last_error()
    q.n_cells[2] = 4
    assert lib.fhdx_params_check(C.byref(q)) == 0
    q = p.copy()
    q.nspecies = 1
    assert lib.fhdx_params_check(C.byref(q)) == 2
    q.nspecies = 9                                         # above MAX_SPECIES = 8
    assert lib.fhdx_params_check(C.byref(q)) == 2
    q = p.copy()
    q.ngc = 1
    assert lib.fhdx_params_check(C.byref(q)) == 2


def test_config_params_pass_check():
    lib = L.lib()
    for cfg in (P.argon_gas(32), P.four_species(16), P.species_prefix(3, 16, 10), P.species_prefix(5, 8, 1),
                P.species_prefix(8, 8, 1), P.no2_n2o4(16),
                P.wall_deck("slip_wall_adiabatic"), P.wall_deck("noslip_reservoir_isothermal"),
                P.wall_deck("temp_grad"), P.wall_deck("RT_3d"), P.wall_deck("giantfluct_3d")):
        assert lib.fhdx_params_check(C.byref(cfg.params)) == 0, cfg.name


def test_wall_helpers_match_the_c_finalize_bit_for_bit():
    """params.set_walls computes the wall mole fractions (GetMolfrac) and the reservoir density
    (GetDensity) with the C library's operation order: fhdx_params_finalize, given only the
    mass fractions and p, T, reproduces them exactly."""
    lib = L.lib()
    p = P.make_params(3, (8, 8, 8), 8e-6, [4.0026, 20.1797, 39.9480], [2.18e-8, 2.58e-8, 3.63e-8])
    P.set_walls(p, [(2, 2), (1, 1), (-1, -1)], [(2, 2), (1, 1), (-1, -1)], [(3, 2), (2, 1), (-1, -1)],
                t_lo=(310.0, 305.0, 0.0), t_hi=(295.0, 0.0, 0.0),
                bc_Yk=[[0.5, 0.3, 0.2], [0.2, 0.3, 0.5], [0.1, 0.6, 0.3], [0] * 3, [0] * 3, [0] * 3],
                p_lo=(1.1e6, 0.0, 0.0))
    q = p.copy()
    for f in range(6):
        for n in range(P.MAX_SPECIES):
            q.bc_Xk[f][n] = 0.0
    q.rho_lo[0] = 0.0
    lib.fhdx_params_finalize(C.byref(q))
    for f in range(3):
        assert list(q.bc_Xk[f]) == list(p.bc_Xk[f]), f
    assert q.rho_lo[0] == p.rho_lo[0] > 0
    assert lib.fhdx_params_check(C.byref(p)) == 0


def test_ambient_pressure_is_the_equation_of_state(orc):
    cfg = P.four_species(4)
    p = cfg.params
    st = orc.HostState(p)
    orc.init_cons(p, st.cu, cfg.rho0, cfg.rhobar, cfg.T_init)
    orc.cons_to_prim(p, st.prim, st.cu)
    g = p.ngc
    assert P.ambient_pressure(p, cfg.rho0, cfg.rhobar, cfg.T_init) == pytest.approx(st.prim[5, g, g, g], rel=1e-14)
