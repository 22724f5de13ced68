"""Writes tests/golden/piso_small.npz: PISO steps of the CPU oracle (the FP64
restatement of NavierStokesSolver, oracle/oracle.cpp) on two small meshes,
used as regression fixtures by tests/test_golden.py.

The reference ships no test vectors for this path, so these are the oracle's
own outputs at the commit that wrote them; they pin the oracle against
unintended change (CPU test) and let the HIP path be checked without running
the oracle (GPU test).  Rerun after an intended change of the oracle:

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path.insert(0, str(REPO))

# name -> (mesh (L, W, H, nx, ny, nz), settings, steps)
CASES = {
    # reference defaults: BiCGSTAB 1e-6 / 1e-8, one corrector, Dirichlet outlet 0
    "kuhn3": ((1.0, 1.0, 1.0, 3, 3, 3),
              dict(dt=0.01, uin=(0.5, 0.0, 0.0), ncorr=1, relax=(1.0, 1.0), pressure_solver=0,
                   p_out=0.0), 3),
    # two correctors, relaxation, Jacobi PCG pressure, outlet pressure 1.5
    "box_ncorr2_cg": ((2.0, 1.0, 1.0, 4, 2, 2),
                      dict(dt=0.01, uin=(0.5, 0.0, 0.0), ncorr=2, relax=(0.5, 0.7),
                           pressure_solver=1, p_out=1.5), 3),
}
RHO, MU = 1060.0, 0.004


def run_oracle(O, mesh, cfg, steps):
    om = O.OracleMesh.structuredRectangularPrism(*mesh)
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    o = O.OracleNS(om, RHO, MU)
    o.setDt(cfg["dt"])
    o.setMeanInletVelocity(cfg["uin"])
    o.setOutlet(0, cfg["p_out"])
    o.setNumberOfCorrections(cfg["ncorr"])
    o.setRelax(*cfg["relax"])
    o.setSolver(1e-6, 1e-8, 1000, cfg["pressure_solver"])
    us, ps, mi, pi = [], [], [], []
    for _ in range(steps):
        if o.pisoStep() != 0:
            raise RuntimeError("oracle step failed")
        st = o.stats()
        us.append(o.getVelocity())
        ps.append(o.getPressure())
        mi.append(st["momentum_iters"])
        pi.append(st["pressure_iters"][:cfg["ncorr"]])
    return (np.array(us), np.array(ps), np.array(mi, np.int32), np.array(pi, np.int32))


def main():
    from oracle import oracle as O
    out = {}
    for name, (mesh, cfg, steps) in CASES.items():
        u, p, mi, pi = run_oracle(O, mesh, cfg, steps)
        out[f"{name}/u"], out[f"{name}/p"] = u, p
        out[f"{name}/momentum_iters"], out[f"{name}/pressure_iters"] = mi, pi
        print(name, u.shape, p.shape, mi.tolist(), pi.tolist())
    np.savez_compressed(HERE / "piso_small.npz", **out)


if __name__ == "__main__":
    main()
