"""fhdx_rk3_step_noalloc: the fast step with its workspace in the library's file-scope device
scratch instead of allocated memory.  Through the C ABI it must equal fhdx_rk3_step bit for bit,
on two calls in a row where the second box is smaller than the first (the scratch keeps the
first call's contents), and it must refuse a box whose workspace exceeds the scratch."""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import perturbed_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def steps(cfg, cu0, nsteps, seed, noalloc):
    import torch

    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    p = cfg.params
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(np.ascontiguousarray(cu0)).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for s in range(nsteps):
        if noalloc:
            L.check(L.lib().fhdx_rk3_step_noalloc(ctx.h, cu.fab, prim.fab, cfg.dt, seed, s, None),
                    "fhdx_rk3_step_noalloc")
        else:
            fc.rk3_step(ctx, cu, prim, cfg.dt, seed, s)
    ctx.check()
    out = cu.t.cpu().numpy(), prim.t.cpu().numpy()
    ctx.close()
    return out


def test_noalloc_equals_allocating_step_twice_second_smaller(orc):
    cases = [P.four_species(20), P.argon_gas(0, shape=(12, 10, 9))]
    for q, cfg in enumerate(cases):
        say(f"case noalloc call {q + 1}: {cfg.name} {tuple(cfg.params.n_cells)}")
        cu0 = perturbed_state(cfg.params, cfg, orc, seed=70 + q).cu
        a = steps(cfg, cu0, 3, 5 + q, noalloc=True)
        b = steps(cfg, cu0, 3, 5 + q, noalloc=False)
        same = np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        say(f"  equals fhdx_rk3_step: {same}")
        assert same


def test_noalloc_refuses_a_box_beyond_the_scratch():
    import torch

    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    cfg = P.four_species(40)
    p = cfg.params
    ctx = L.Context(p)
    cu = fc.MultiFab.cell(p, p.nvars)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    rc = L.lib().fhdx_rk3_step_noalloc(ctx.h, cu.fab, prim.fab, cfg.dt, 1, 0, None)
    msg = L.lib().fhdx_last_error().decode()
    say(f"case noalloc 40^3 four species: status {rc}, '{msg}'")
    assert rc == 2 and "MiB" in msg
    torch.cuda.synchronize()
    ctx.close()
