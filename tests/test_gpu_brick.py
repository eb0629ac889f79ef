"""cmdg_create_brick (a brick given by its break points) against cmdg_create on the same
brick's per-node vgeo/sgeo/vmaps: the same element records, static classes and
traversal order, so right-hand sides, steps and diagnostics agree bit for bit."""
import math

import numpy as np
import pytest

from cmdg import abi, atmos, closure
from cmdg.grid import uniform_brick

pytestmark = pytest.mark.gpu


def _cases():
    prm = atmos.ideal_gas_parameters()
    tgv = dict(n=(6, 5, 4), lo=(0, 0, 0), hi=(2 * math.pi, 2 * math.pi, 2 * math.pi), periodicity=(True,) * 3,
               boundary=((1, 1), (1, 1), (1, 2)))
    les = dict(n=(4, 3, 5), lo=(0, 0, 0), hi=(400.0, 300.0, 500.0), periodicity=(True, True, False),
               boundary=((0, 0), (0, 0), (1, 2)))
    yield "tgv_fp64", prm, tgv, atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, C_smag=0.21), abi.FP64
    yield "tgv_learned", prm, tgv, atmos.AtmosConfig(prm, turbulence=abi.LEARNED_CLOSURE, C_smag=0.21,
                                                     viscosity=0.02), abi.MIXED
    lk = dict(orientation=abi.FLAT_ORIENTATION, ref_profile=atmos.DryAdiabaticProfile(300.0, 0.0),
              gravity_source=True, C_smag=0.23)
    yield "les_fp64", prm, les, atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, **lk), abi.FP64
    yield "les_mixed", prm, les, atmos.AtmosConfig(prm, turbulence=abi.SMAGORINSKY_LILLY, **lk), abi.MIXED


@pytest.mark.parametrize("case", [c[0] for c in _cases()])
def test_brick_context_is_bitwise_the_array_context(case):
    import torch
    from cmdg.dg import DGModel, LSRK54CarpenterKennedy, solve, to_device
    print(f"[brick] {case}", flush=True)
    name, prm, g, cfg, precision = next(c for c in _cases() if c[0] == case)
    full = uniform_brick(g["n"], g["lo"], g["hi"], N=4, periodicity=g["periodicity"], boundary=g["boundary"])
    lean = uniform_brick(g["n"], g["lo"], g["hi"], N=4, periodicity=g["periodicity"], boundary=g["boundary"],
                         lean=True)
    aux = atmos.init_aux(cfg, full)
    if name.startswith("tgv"):
        Q0 = atmos.taylor_green(prm, full, mach=0.1)
    else:
        Q0 = atmos.rising_bubble(prm, full, radius=100.0, centre=(200.0, 150.0, 150.0), seed=5)
    out = []
    for grid in (full, lean):
        dg = DGModel(cfg, grid, aux=aux, precision=precision)
        if cfg.turbulence == abi.LEARNED_CLOSURE:
            dg.set_closure(closure.MLPParams.constant(0.21 ** 2))
        Q = to_device(Q0)
        T = torch.full_like(Q, float("nan"))
        dg(T, Q, None, 0.0)
        solver = LSRK54CarpenterKennedy(dg, Q, dt=2e-4 if name.startswith("tgv") else 5e-3, t0=0.0)
        solve(Q, solver, numberofsteps=3)
        out.append((dg.geometry_kind, dg.static_aux_classes, dg.tiled_traversal, T.cpu(), Q.cpu(),
                    [dg.weightedsum(Q, s) for s in range(5)]))
        dg.close()
    a, b = out
    assert a[:3] == b[:3]
    assert torch.equal(a[3], b[3]) and torch.equal(a[4], b[4])
    assert a[5] == b[5]
