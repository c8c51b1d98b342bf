"""The full-rank FP64 plume solver on the device (csrc/fluid.hip: FLUID_3D_MIC
addSmokeColumn + step with the appendStreams snapshots) against the FP64
oracle (oracle/fluid_oracle.py), each step from the device's own previous
state.  Both are FP64; they differ in the summation order of the CG's dot
products only, so the fields agree to 1e-8 relative (the CG stops at a
residual maximum of 1e-10)."""
import numpy as np
import pytest

from cfs_plume_amd import fluid as FL
from oracle import fluid_oracle as FO

pytestmark = pytest.mark.gpu

FULL = (12, 24, 12)
TOL = 1e-8


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("vorticity", [0.0, 2.0])
def test_fluid_steps_match_the_oracle(gpu, vorticity):
    print(f"[case] full-rank plume {FULL} vorticity={vorticity}", flush=True)
    f = FL.FluidSolver(FULL, vorticity=vorticity)
    o = FO.Solver(FULL, FO.default_params(FULL, vorticity))
    steps = 4
    snaps = f.new_snapshots(steps)
    for t in range(steps):
        v0, d0, h0 = f.velocity.numpy(), f.density.numpy(), f.heat.numpy()
        f.add_smoke()
        f.step(snaps, t)
        its_gpu, max_gpu = f.last_solve()
        d1, h1 = o.add_smoke(d0, h0)
        v, d, h, sn, (its, mx) = o.step(v0, d1, h1)
        errs = {"velocity": rel(f.velocity.numpy(), v), "density": rel(f.density.numpy(), d),
                "heat": rel(f.heat.numpy(), h)}
        for name in FL.STREAMS:
            errs[name] = rel(snaps.numpy(name)[t], sn[name])
        worst = max(errs, key=errs.get)
        print(f"[fluid] step {t}: CG {its_gpu} vs {its} iterations, residual max {max_gpu:.2e} vs {mx:.2e}; "
              f"worst {worst} {errs[worst]:.2e}", flush=True)
        assert abs(its_gpu - its) <= 1
        assert max_gpu <= 1e-10 or its_gpu == f.params.iterations
        assert max(errs.values()) <= TOL, errs
    assert np.linalg.norm(f.velocity.numpy()) > 0.0
    f.close()


def test_fluid_rest_and_parameters(gpu):
    p = FL.default_params((48, 64, 48))
    assert p.dt == 0.1 and p.dx == 1.0 / 64 and p.vorticity_eps == 0.0 and p.iterations == 1000
    assert list(p.boundaries) == [1, 1, 1, 1, 0, 0]
    f = FL.FluidSolver((8, 10, 8))
    f.step()
    assert not np.any(f.velocity.numpy()) and not np.any(f.pressure.numpy())
    assert f.last_solve()[0] == 0
    f.close()
