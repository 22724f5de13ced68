"""Regression fixtures: tests/golden/piso_small.npz holds PISO steps of the CPU
oracle on two small meshes (written by tests/golden/make_golden.py; the
reference ships no test vectors for this path).  The oracle must still
reproduce them bit for bit, and the HIP path must match them without the
oracle running: equal iteration counts and relative L2 <= 1e-8 per step (in
practice the same bits)."""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
_spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN / "piso_small.npz") as z:
        return {k: z[k] for k in z.files}


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("name", list(make_golden.CASES))
def test_oracle_reproduces_golden(O, gold, name):
    mesh, cfg, steps = make_golden.CASES[name]
    u, p, mi, pi = make_golden.run_oracle(O, mesh, cfg, steps)
    assert np.array_equal(u, gold[f"{name}/u"])
    assert np.array_equal(p, gold[f"{name}/p"])
    assert np.array_equal(mi, gold[f"{name}/momentum_iters"])
    assert np.array_equal(pi, gold[f"{name}/pressure_iters"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(make_golden.CASES))
def test_gpu_matches_golden(P, gold, name):
    mesh, cfg, steps = make_golden.CASES[name]
    pm = P.Mesh.structuredRectangularPrism(*mesh)
    pm.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    g = P.NavierStokesSolver()
    g.setOutletType(0)
    g.initialize(pm, make_golden.RHO, make_golden.MU)
    g.setDt(cfg["dt"])
    g.setMeanInletVelocity(cfg["uin"])
    g.setOutletPressure(cfg["p_out"])
    g.setNumberOfCorrections(cfg["ncorr"])
    g.setRelaxU(cfg["relax"][0])
    g.setRelaxP(cfg["relax"][1])
    g.setOptions(momentum_tol=1e-6, pressure_tol=1e-8, max_iter=1000,
                 pressure_solver=cfg["pressure_solver"])
    for s in range(steps):
        assert g.pisoStepStatus() == 0, (name, s, P.last_error())
        st = g.stepStats()
        assert st["momentum_iters"] == gold[f"{name}/momentum_iters"][s].tolist(), (name, s)
        assert st["pressure_iters"] == gold[f"{name}/pressure_iters"][s].tolist(), (name, s)
        assert relerr(g.getVelocity(), gold[f"{name}/u"][s]) <= 1e-8, (name, s)
        assert relerr(g.getPressure(), gold[f"{name}/p"][s]) <= 1e-8, (name, s)
