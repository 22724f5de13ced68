#!/usr/bin/env python3
"""PISO steps/s on one MI355X: mixed-precision vs FP64, both at 1e-8 relative
residual, on the 96^3 unit-cube Kuhn channel (5,308,416 tets, 21,567,171 P2
velocity DOFs, 912,673 pressure DOFs).

One step = one NavierStokesSolver::pisoStep (convection, momentum right-hand
side, batched 3-component momentum solve, divergence, pressure solve,
velocity correction, boundary resets, NaN check) on device-resident data.
Prints ONE JSON line.  The CPU oracle is timed beside it on a bounded sample
(see cpu_baseline.sample).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "cell-fluid-structure-analysis_amd"))
sys.path.insert(0, str(REPO))

HBM_PEAK_GBS = 8000.0
# rocprofv3 FETCH_SIZE + WRITE_SIZE of the FP32 momentum SpMV at 96^3, scaled
# by the counters' readings for a plain 1 GiB copy (summary and commands in
# this file)
TRAFFIC_PROFILE = REPO / "profiles" / "k_csr_stream_f32_momentum_kuhn96_traffic.md"
ELEM_TRAFFIC_PROFILE = REPO / "profiles" / "k_elem_spmv_momentum_kuhn96_traffic.md"


def profiled_traffic(path=TRAFFIC_PROFILE):
    try:
        for line in path.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None
METRIC = "PISO steps/s, mixed vs FP64 GPU at 1e-8 rel. residual; % HBM roofline"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, w, k=3):
    """Algorithmic bytes of one CSR SpMV with k right-hand sides sharing the
    matrix: values + int32 columns + int32 row pointers + read x, write y."""
    return (w + 4) * nnz + 4 * (n + 1) + 2 * k * w * n


def elem_bytes(info, nu, nt):
    """Algorithmic bytes of one element-patch SpMV (FP32, 3 right-hand sides):
    the visit records (32 B each), the per-tet geometry (8 floats: rho V / dt
    and the 3 x 3 Gram matrix of the scaled gradients), the patches' node
    tables (int32) and slot starts (uint16), the row penalties, read x and
    write y once (12 B per node each); the fused inner product's operand is
    not counted, as for the CSR kernel."""
    return (32 * info["visits"] + 32 * nt + 4 * info["node_entries"]
            + 2 * (nu + info["patches"]) + 4 * nu + 2 * 3 * 4 * nu)


def make_solver(P, mesh, prof, args, precision, momentum_solver=0, initial_guess=0):
    g = P.NavierStokesSolver()
    g.initialize(mesh, args.rho, args.mu)
    g.setDt(args.dt)
    g.setNumberOfCorrections(args.corrections)
    g.setInletProfile(prof)
    g.setOptions(momentum_tol=args.momentum_tol, pressure_tol=args.pressure_tol,
                 precision=precision, momentum_solver=momentum_solver, timing=2,
                 check_every=args.check_every, stop_rule=args.stop_rule,
                 pressure_solver=args.pressure_solver, inner_tol=args.inner_tol,
                 initial_guess=initial_guess,
                 reuse_pressure_correction=0 if args.resolve_correctors else 1,
                 elem_spmv=args.elem_spmv if precision else 0)
    return g


def run(P, g, steps, warmup, torch):
    for _ in range(warmup):
        if not g.pisoStep():
            raise RuntimeError(f"warmup step failed: {P.last_error()}")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        if not g.pisoStep():
            raise RuntimeError(f"step failed: {P.last_error()}")
        stats.append(g.stepStats())
    torch.cuda.synchronize()
    return time.perf_counter() - t0, stats


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_whole_steps(P, O, workload, steps):
    """Whole oracle pisoSteps (convection, right-hand sides, Eigen-form Krylov
    solves, divergence, correction: everything the reference's step does), one
    core, timed after initialize, on the cylinder workloads: config 1 (length
    10, radius 1, 16 divisions, max element 0.1; CG pressure) and config 2 (the
    README cylinder, 2 correctors).  The first step starts from rest."""
    if workload == "config1":
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
        R, ncorr, psolver = 1.0, 1, 1
    else:
        pm = P.Mesh.cylinder(30.0, 2.0, 35, 0.03)
        R, ncorr, psolver = 2.0, 2, 0
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    t0 = time.perf_counter()
    o = O.OracleNS(om, 1060.0, 0.004)
    t_init = time.perf_counter() - t0
    X = np.vstack([om.vertices(), om.edgeNodes()])
    g = np.zeros((len(X), 3))
    g[:, 0] = 0.5 * np.clip(1.0 - (X[:, 1] ** 2 + X[:, 2] ** 2) / R ** 2, 0.0, None)
    o.setDt(0.01)
    o.setNumberOfCorrections(ncorr)
    o.setInletProfile(g)
    o.setSolver(1e-6, 1e-8, 1000, psolver)
    t0 = time.perf_counter()
    for _ in range(steps):
        if o.pisoStep() != 0:
            raise RuntimeError("oracle step failed")
    t = time.perf_counter() - t0
    sz = o.sizes()
    return {"steps_per_s": steps / t, "steps": steps, "s_per_step": t / steps,
            "initialize_s": t_init, "tets": sz["nt"], "velocity_dofs": 3 * sz["nu"],
            "pressure_dofs": sz["nv"], "last_iters": o.stats()}


def jacobi_pressure_iters(g, args):
    """Jacobi-CG iterations of the last step's pressure system (rhs = -div(u*)/dt,
    the outlet pressure being 0) on the GPU: what the oracle's CG would take,
    when the timed steps used the AMG-preconditioned solve."""
    rhs = -g.computeDivergence(g.getVelocityStar()) / args.dt
    _, r = g.krylovSolve("Lbc", rhs, method=1, tol=args.pressure_tol, maxit=20000)
    return r["iters"][0]


def cpu_baseline(O, g, fp64_stats, sz, pressure_method, pressure_iters=None):
    """Eigen-form BiCGSTAB / Jacobi CG of the oracle (one core, inner products
    in Eigen's packet order) timed for a few iterations on the workload's own
    momentum and pressure operators, then scaled by the iteration counts of
    the timed FP64 GPU steps (the oracle takes the same counts: see
    tests/test_gpu_parity.py; with the AMG pressure solve on the GPU, the
    pressure count is the Jacobi-CG count of the same system, pressure_iters).
    An extrapolation: a whole oracle step at 96^3 takes minutes; whole steps
    are timed on the cylinders (cpu_whole_steps)."""
    O.set_dot_mode(1)
    try:
        return _cpu_baseline(O, g, fp64_stats, sz, pressure_method, pressure_iters)
    finally:
        O.set_dot_mode(0)


def _cpu_baseline(O, g, fp64_stats, sz, pressure_method, pressure_iters=None):
    rp, col, val = g.getMatrix("lhs")
    rng = np.random.default_rng(0)
    k_mom = 30
    b = rng.uniform(-1, 1, sz["nu"])
    t0 = time.perf_counter()
    O.csr_krylov(rp, col, val, b, method=0, tol=0.0, maxit=k_mom)
    t_mom = (time.perf_counter() - t0) / k_mom
    del rp, col, val
    rp, col, val = g.getMatrix("Lbc")
    bp = rng.uniform(-1, 1, sz["nv"])
    k_p = 400
    t0 = time.perf_counter()
    O.csr_krylov(rp, col, val, bp, method=pressure_method, tol=0.0, maxit=k_p)
    t_p = (time.perf_counter() - t0) / k_p
    mom = np.mean([sum(s["momentum_iters"]) for s in fp64_stats])
    if pressure_iters is None:
        pre = np.mean([sum(s["pressure_iters"]) for s in fp64_stats])
    else:
        pre = float(np.mean([len(s["pressure_iters"]) for s in fp64_stats]) * pressure_iters)
    t_step = mom * t_mom + pre * t_p
    return {"value": 1.0 / t_step, "unit": "PISO steps/s", "cores": 1, "kind": "port",
            "method": "extrapolated from timed Krylov iterations",
            "sample": (f"oracle Eigen-form BiCGSTAB, {k_mom} iterations on the {sz['nu']}-row "
                       f"momentum matrix ({t_mom:.3f} s/iter per component) and {k_p} "
                       f"{'Jacobi-CG' if pressure_method == 1 else 'BiCGSTAB'} on the "
                       f"{sz['nv']}-row pressure matrix ({t_p * 1e3:.2f} ms/iter), scaled by the "
                       f"FP64 run's mean {mom:.1f} momentum + {pre:.1f} pressure iterations per "
                       f"step" + (" (the pressure count: Jacobi-CG on the last step's pressure "
                                  "system, the GPU steps having used the AMG-preconditioned "
                                  "solve)" if pressure_iters is not None else "") +
                       "; element kernels (convection, divergence, right-hand sides) "
                       "not counted, which favours the CPU"),
            "cpu": cpu_model(), "nproc": os.cpu_count()}


PRESSURE_SOLVERS = {0: "bicgstab_jacobi", 1: "cg_jacobi", 2: "cg_amg"}

# FP64 variants: (name, momentum solver); the reference's BiCGSTAB first
FP64 = [("fp64", 0), ("fp64_cg", 1), ("fp64_cgpoly", 2), ("fp64_cgpoly2", 3)]
MOMENTUM_SOLVERS = {0: "bicgstab_jacobi", 1: "cg_jacobi", 2: "cg_chebyshev_poly",
                    3: "cg_chebyshev_poly2"}
# mixed-precision variants: (name, precision, momentum solver, initial guess)
#   precision 1: FP32 inner work for momentum and pressure (with the AMG pressure
#   solve: an FP32 V-cycle inside the FP64 CG), 2: momentum only
#   momentum solver 0: BiCGSTAB (the reference's), 1: Jacobi CG (lhs is SPD),
#   2: CG with the degree-1 Chebyshev polynomial preconditioner in D^-1 A, 3: degree 2
MIXED = [("mixed_bicgstab_all", 1, 0, 0), ("mixed_bicgstab", 2, 0, 0),
         ("mixed_cg_all", 1, 1, 0), ("mixed_cg", 2, 1, 0), ("mixed_cg_warm", 2, 1, 1),
         ("mixed_cgpoly_all", 1, 2, 0), ("mixed_cgpoly", 2, 2, 0),
         ("mixed_cgpoly2_all", 1, 3, 0), ("mixed_cgpoly2", 2, 3, 0)]
# the FP64 variant with the same momentum solver as each mixed variant
SAME_SOLVER = {"mixed_bicgstab_all": "fp64", "mixed_bicgstab": "fp64", "mixed_cg_all": "fp64_cg",
               "mixed_cg": "fp64_cg", "mixed_cg_warm": "fp64_cg",
               "mixed_cgpoly_all": "fp64_cgpoly", "mixed_cgpoly": "fp64_cgpoly",
               "mixed_cgpoly2_all": "fp64_cgpoly2", "mixed_cgpoly2": "fp64_cgpoly2"}
# variants that keep the reference's zero initial guess compete for the headline
HEADLINE = {"mixed_bicgstab_all", "mixed_bicgstab", "mixed_cg_all", "mixed_cg",
            "mixed_cgpoly_all", "mixed_cgpoly", "mixed_cgpoly2_all", "mixed_cgpoly2"}
VEL_ERR_MAX = 1e-7


def cylinder_mesh(P):
    """The README cylinder (length 30, radius 2, 35 divisions, max element 0.03)
    from the library's own mesher; x is the flow axis."""
    mesh = P.Mesh.cylinder(30.0, 2.0, 35, 0.03)
    mesh.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    return mesh


def fourier_inlet_basis(X, R):
    """u_x basis on the inlet disc: (1 - r^2/R^2) times 1, cos t, sin t, cos 2t."""
    y, z = X[:, 1], X[:, 2]
    base = np.clip(1.0 - (y * y + z * z) / R ** 2, 0.0, None)
    th = np.arctan2(z, y)
    B = np.zeros((4, len(X), 3))
    for k, f in enumerate([np.ones_like(th), np.cos(th), np.sin(th), np.cos(2 * th)]):
        B[k, :, 0] = base * f
    return B


def adjoint_workload(P, args, torch):
    """Gradient of J = 1/2 sum_Outlet |u|^2 after `steps` PISO steps from rest on the
    cylinder, w.r.t. 4 inlet Fourier coefficients c ~ 0.5 + N(0, 0.05)
    (numpy default_rng(0)) and the viscosity, by the library's hand-written
    adjoint (per-step checkpoints, u* and p' recomputed); every component is
    checked against a central difference with h = 1e-6 of the GPU forward."""
    R = 2.0
    t0 = time.perf_counter()
    mesh = cylinder_mesh(P)
    X = np.vstack([mesh.vertices(), mesh.edgeNodes()])
    B = fourier_inlet_basis(X, R)
    c = 0.5 + np.random.default_rng(0).normal(0.0, 0.05, 4)
    dt = 0.01 if args.dt is None else args.dt
    ncorr = 2 if args.corrections is None else args.corrections
    tol = args.adjoint_tol
    g = P.NavierStokesSolver()
    g.initialize(mesh, args.rho, args.mu)
    g.setDt(dt)
    g.setNumberOfCorrections(ncorr)
    g.setOptions(momentum_tol=tol, pressure_tol=tol, stop_rule=1, max_iter=20000,
                 pressure_solver=args.pressure_solver)
    sz = g.sizes()
    log(f"adjoint workload: mesh {mesh.sizes()}, solver ready in {time.perf_counter() - t0:.1f}s")
    lab = np.r_[mesh.p1BC(), mesh.p2BC()]
    outlet = lab == 2

    def J_of(coef):
        u = g.adjointForward(args.steps, np.tensordot(coef, B, axes=1))
        return 0.5 * float((u[outlet] ** 2).sum())

    g.adjointGradient(args.steps, B, c)  # warmup
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        J, dc, dmu = g.adjointGradient(args.steps, B, c)
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) / reps
    log(f"adjoint gradient {1e3 * t_grad:.1f} ms")
    t0 = time.perf_counter()
    J_fwd = J_of(c)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    h = 1e-6
    fd = []
    for k in range(4):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        fd.append((J_of(cp) - J_of(cm)) / (2 * h))
        log(f"fd c{k}: {fd[-1]:.9e} vs adjoint {dc[k]:.9e}")
    g.setViscosity(args.mu + h)
    Jp = J_of(c)
    g.setViscosity(args.mu - h)
    Jm = J_of(c)
    g.setViscosity(args.mu)
    fdmu = (Jp - Jm) / (2 * h)
    # five-point stencil with H = 10 h: truncation O(H^4) (about 4e-11 relative
    # here) and 10x less solver noise than the h = 1e-6 quotient, whose signal
    # J(mu + h) - J(mu - h) is only ~1e-7 of J
    H = 10 * h
    Jv = {}
    for m in (-2, -1, 1, 2):
        g.setViscosity(args.mu + m * H)
        Jv[m] = J_of(c)
    g.setViscosity(args.mu)
    fdmu5 = (Jv[-2] - 8 * Jv[-1] + 8 * Jv[1] - Jv[2]) / (12 * H)
    log(f"fd mu: central h {fdmu:.9e}, five-point 10h {fdmu5:.9e}, adjoint {dmu:.9e}")
    scale = float(np.abs(dc).max())
    err_c = [abs(f - d) / scale for f, d in zip(fd, dc)]
    err_mu = abs(fdmu - dmu) / abs(dmu)
    log(f"J {J:.12e} (forward only {J_fwd:.12e}); dJ/dc {dc}; fd {fd}; dJ/dmu {dmu:.6e} fd {fdmu:.6e}")
    out = {"metric": "adjoint gradient of the outlet kinetic energy (4 inlet coefficients + "
                     "viscosity), checked against central differences",
           "value": 1.0 / t_grad, "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 1, "ms_per_gradient": 1e3 * t_grad, "ms_forward_only": 1e3 * t_fwd,
           "higher_is_better": True, "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "adjoint_cylinder_L30_R2_div35_a0.03",
                      "tets": sz["nt"], "velocity_dofs": 3 * sz["nu"], "pressure_dofs": sz["nv"],
                      "piso_steps": args.steps, "corrections": ncorr, "dt": dt,
                      "density": args.rho, "viscosity": args.mu, "coefficients": c.tolist(),
                      "solver_tol": tol, "stop_rule": "jacobi_scaled", "fd_step": h},
           "J": J, "J_forward_only": J_fwd, "dJ_dcoef": dc.tolist(), "dJ_dviscosity": dmu,
           "fd_dJ_dcoef": fd, "fd_dJ_dviscosity": fdmu, "fd5_dJ_dviscosity_10h": fdmu5,
           "fd5_error_viscosity_rel": abs(fdmu5 - dmu) / abs(dmu),
           "fd_error_coef_normwise": err_c, "fd_error_viscosity_rel": err_mu,
           "fd_bound": 1e-5, "fd_ok": bool(max(err_c) <= 1e-5 and err_mu <= 1e-5)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="20 (kuhn) / 50 (cyl2) / 10 (adjoint)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=96, help="cells per side of the unit cube")
    ap.add_argument("--workload", choices=["kuhn", "cyl2", "adjoint"], default="kuhn",
                    help="kuhn: unit-cube Kuhn channel (default); cyl2: the README cylinder "
                         "(length 30, radius 2, 35 divisions, max element 0.03; parabolic "
                         "inlet peak 0.5, dt 0.01, 2 correctors, 50 steps); adjoint: "
                         "gradient of the outlet kinetic energy after 10 PISO steps on the "
                         "cylinder w.r.t. 4 inlet Fourier coefficients and the viscosity, "
                         "checked against central differences")
    ap.add_argument("--rho", type=float, default=1060.0)
    ap.add_argument("--mu", type=float, default=0.004)
    ap.add_argument("--re", type=float, default=100.0)
    ap.add_argument("--dt", type=float, default=None, help="1e-3 (kuhn) / 0.01 (cyl2)")
    ap.add_argument("--corrections", type=int, default=None, help="1 (kuhn) / 2 (cyl2)")
    ap.add_argument("--pressure-tol", type=float, default=1e-8)
    ap.add_argument("--momentum-tol", type=float, default=1e-10,
                    help="tight enough that two correct solves agree to 1e-7 in velocity")
    ap.add_argument("--check-every", type=int, default=8)
    ap.add_argument("--inner-tol", type=float, default=1e-4,
                    help="relative tolerance of each FP32 inner solve (mixed variants)")
    ap.add_argument("--stop-rule", type=int, default=1,
                    help="0: |r| <= tol |b| (reference; its 1e10 inlet-penalty rows dominate "
                         "|b|, so the interior is not solved to tol), 1: Jacobi-scaled")
    ap.add_argument("--pressure-solver", type=int, default=2, choices=[0, 1, 2],
                    help="0 BiCGSTAB, 1 Jacobi CG, 2 AMG-preconditioned CG")
    ap.add_argument("--adjoint-tol", type=float, default=1e-14,
                    help="Krylov tolerance (Jacobi-scaled rule) of the adjoint workload")
    ap.add_argument("--resolve-correctors", action="store_true",
                    help="re-solve the pressure system in every PISO corrector as the reference "
                         "does, instead of reusing the first corrector's p' (same bits; "
                         "matters only with 2 or more correctors)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-cpu-steps", action="store_true",
                    help="skip the whole-step oracle timing on the two cylinders (~1 min)")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--only", default=None,
                    help="run only these variants, comma-separated (fp64 always runs)")
    ap.add_argument("--elem-spmv", type=int, default=1, choices=[0, 1],
                    help="mixed variants: FP32 momentum operator applied tet by tet (1) or "
                         "from its assembled CSR copy (0)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the PISO "
              f"path is built and measured on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import torch
    import cfsa_ns as P

    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 4,
                          "copy_probe_bytes_written": dst.numel() * 4, "launches": 5}))
        return

    if args.steps is None:
        args.steps = {"kuhn": 20, "cyl2": 50, "adjoint": 10}[args.workload]
    if args.workload == "adjoint":
        adjoint_workload(P, args, torch)
        return

    t0 = time.perf_counter()
    n = args.n
    cyl = args.workload == "cyl2"
    if args.dt is None:
        args.dt = 0.01 if cyl else 1e-3
    if args.corrections is None:
        args.corrections = 2 if cyl else 1
    if cyl:
        R = 2.0
        mesh = cylinder_mesh(P)
        U = 0.5
        X = np.vstack([mesh.vertices(), mesh.edgeNodes()])
        prof = np.zeros((len(X), 3))
        prof[:, 0] = U * np.clip(1.0 - (X[:, 1] ** 2 + X[:, 2] ** 2) / R ** 2, 0.0, None)
        workload = "cylinder_L30_R2_div35_a0.03"
        mesh_desc = "own cylinder mesher (TetGen is not available), length 30, radius 2"
    else:
        mesh = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
        mesh.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
        U = args.re * args.mu / (args.rho * 1.0)
        X = np.vstack([mesh.vertices(), mesh.edgeNodes()])
        prof = np.zeros((len(X), 3))
        prof[:, 0] = U * 16.0 * X[:, 1] * (1 - X[:, 1]) * X[:, 2] * (1 - X[:, 2])
        workload = f"kuhn{n}_channel"
        mesh_desc = f"unit cube {n}^3 Kuhn 6-tet"
    log(f"mesh {workload}: {mesh.sizes()} in {time.perf_counter() - t0:.1f}s")

    def spmv_avg(st):
        return sum(s["spmv_ms"] for s in st) / max(1, sum(s["spmv_launches"] for s in st))

    res = {}
    for name, msolver in FP64:
        if args.only and name != "fp64" and name not in args.only.split(","):
            continue
        t0 = time.perf_counter()
        g = make_solver(P, mesh, prof, args, 0, msolver)
        sz = g.sizes()
        log(f"{name} solver ready in {time.perf_counter() - t0:.1f}s: {sz}")
        el, st = run(P, g, args.steps, args.warmup, torch)
        log(f"{name}: {args.steps} steps in {el:.3f}s; last {st[-1]}")
        res[name] = {"steps_per_s": args.steps / el, "ms_per_step": 1e3 * el / args.steps,
                     "momentum_solver": MOMENTUM_SOLVERS[msolver],
                     "momentum_iters": st[-1]["momentum_iters"],
                     "pressure_iters": st[-1]["pressure_iters"], "spmv_avg_ms": spmv_avg(st)}
        if name == "fp64":
            g64, el64, st64 = g, el, st
        else:
            del g
    u64, p64 = g64.getVelocity(), g64.getPressure()
    d64 = g64.computeDivergence(u64)
    # a caller that reads the fields back to the host after every step (the
    # reference's getVelocity / getPressure return host vectors) also pays
    # this: pageable copies, and the library's snapshot + pinned copy started
    # right after the step (no later work to overlap: the worst case)
    t0 = time.perf_counter()
    for _ in range(3):
        g64.getVelocity()
        g64.getPressure()
    host_copy_pageable_ms = 1e3 * (time.perf_counter() - t0) / 3
    pu, pp = P.PinnedArray((sz["nu"], 3)), P.PinnedArray((sz["nv"],))
    g64.getFieldsAsync(pu.array, pp.array)
    g64.waitFields()
    t0 = time.perf_counter()
    for _ in range(3):
        g64.getFieldsAsync(pu.array, pp.array)
        g64.waitFields()
    host_copy_ms = 1e3 * (time.perf_counter() - t0) / 3
    assert np.array_equal(pu.array, u64) and np.array_equal(pp.array, p64)
    del pu, pp
    b64 = spmv_bytes(sz["nu"], sz["nnzA"], 8)
    b32 = spmv_bytes(sz["nu"], sz["nnzA"], 4)
    primary, best = "fp64", None
    if not args.fp64_only:
        for name, precision, msolver, guess in MIXED:
            if args.only and name not in args.only.split(","):
                continue
            t0 = time.perf_counter()
            g = make_solver(P, mesh, prof, args, precision, msolver, guess)
            el, st = run(P, g, args.steps, args.warmup, torch)
            log(f"{name}: {args.steps} steps in {el:.3f}s; last {st[-1]}")
            einfo = g.elemInfo()
            u, p = g.getVelocity(), g.getPressure()
            d = g.computeDivergence(u)
            err = {"velocity_rel_l2": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
                   "pressure_rel_l2": float(np.linalg.norm(p - p64) / np.linalg.norm(p64)),
                   "divergence_rel_l2": float(np.linalg.norm(d - d64) / np.linalg.norm(d64))}
            res[name] = {"steps_per_s": args.steps / el, "ms_per_step": 1e3 * el / args.steps,
                         "precision": precision,
                         "momentum_solver": MOMENTUM_SOLVERS[msolver],
                         "initial_guess": "u^n" if guess else "zero",
                         "momentum_inner_iters": st[-1]["momentum_iters"],
                         "momentum_refinements": st[-1]["momentum_refinements"],
                         "pressure_iters": st[-1]["pressure_iters"],
                         "pressure_refinements": st[-1]["pressure_refinements"],
                         "spmv_avg_ms": spmv_avg(st),
                         "speedup_vs_fp64_same_solver": (
                             res[SAME_SOLVER[name]]["ms_per_step"] / (1e3 * el / args.steps)
                             if SAME_SOLVER[name] in res else None),
                         "speedup_vs_fp64_bicgstab": el64 / el, "error_vs_fp64": err,
                         "momentum_operator_fp32": "element patches" if einfo["built"]
                         else "assembled CSR copy"}
            if einfo["built"]:
                res[name]["elem_patches"] = einfo
            ok = err["velocity_rel_l2"] <= VEL_ERR_MAX and name in HEADLINE
            if ok and (best is None or el < best[1]):
                best = (name, el)
            del g
        if best is not None:
            primary = best[0]
    if cyl:
        # io::exportToVtk of the fluid fields (VtkExport.cpp:238-268), the
        # runner's output path, on the README cylinder: host timing of one write
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            t0 = time.perf_counter()
            g64.exportVtk(os.path.join(td, "fluid.vtk"))
            res["vtk_write_s"] = time.perf_counter() - t0
            res["vtk_bytes"] = os.path.getsize(os.path.join(td, "fluid.vtk"))
        log(f"vtk write {res['vtk_write_s']:.2f}s, {res['vtk_bytes']} bytes")
    # value-only re-assembly of M, K, L, G into the fixed patterns (what a coupled
    # FSI step pays after the mesh moves), host-timed including the sync
    g64.build_sparse_matrices()
    t0 = time.perf_counter()
    for _ in range(3):
        g64.build_sparse_matrices()
    res["reassembly_ms"] = 1e3 * (time.perf_counter() - t0) / 3
    # both Krylov SpMV layouts on the momentum operator (3 right-hand sides)
    lay = {}
    for layout, name in ((1, "csr_stream"), (2, "sliced_ell")):
        g64.setOptions(spmv_layout=layout)
        t64 = g64.timeMomentumSpmv(20, 0)
        t32 = g64.timeMomentumSpmv(20, 1)
        lay[name] = {"fp64_ms": t64, "fp64_GB/s": b64 / t64 / 1e6, "fp32_ms": t32,
                     "fp32_GB/s": b32 / t32 / 1e6}
    g64.setOptions(spmv_layout=0)
    res["spmv_layouts"] = lay
    log(f"spmv layouts: {lay}")

    if primary == "fp64":
        ms = res["fp64"]["spmv_avg_ms"]
        roof = {"bound": "hbm", "kernel": "k_csr_stream<double,3> (momentum SpMV, 3 rhs, "
                "fused inner products)", "bytes_per_launch": b64, "avg_launch_ms": ms}
        ach = b64 / ms / 1e6
    elif "elem_patches" in res[primary]:
        ms = res[primary]["spmv_avg_ms"]
        be = elem_bytes(res[primary]["elem_patches"], sz["nu"], sz["nt"])
        roof = {"bound": "hbm", "kernel": "k_elem_spmv<OpAxDot> (FP32 momentum operator tet by "
                "tet from element patches, 3 rhs, fused inner products)",
                "bytes_per_launch": be, "avg_launch_ms": ms}
        ach = be / ms / 1e6
    else:
        ms = res[primary]["spmv_avg_ms"]
        roof = {"bound": "hbm", "kernel": "k_csr_stream<float,3> (momentum SpMV, 3 rhs, "
                "fused inner products)", "bytes_per_launch": b32, "avg_launch_ms": ms}
        ach = b32 / ms / 1e6
    tpath = ELEM_TRAFFIC_PROFILE if "k_elem_spmv" in roof["kernel"] else TRAFFIC_PROFILE
    traffic = profiled_traffic(tpath) if (primary != "fp64" and not cyl and n == 96) else None
    roof.update({"achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                 "traffic_source": str(tpath.relative_to(REPO)) if traffic else None})
    if "k_elem_spmv" in roof["kernel"]:
        # the assembled operator's FP32 CSR-stream product on the same system,
        # timed alone (20 launches), for comparison
        t32 = lay["csr_stream"]["fp32_ms"]
        res["roofline_csr_fp32"] = {"kernel": "k_csr_stream<float,3>", "bytes_per_launch": b32,
                                    "avg_launch_ms_alone": t32,
                                    "frac": b32 / t32 / 1e6 / HBM_PEAK_GBS}
        # the same product as the CSR kernel delivers it, per launch of the
        # element kernel: the CSR form's algorithmic bytes over the element
        # kernel's time (a rate no CSR kernel can exceed without reaching
        # this fraction of the HBM peak); not the element kernel's roofline
        res["momentum_product_csr_equivalent"] = {
            "csr_bytes_per_launch": b32, "elem_ms": ms,
            "GB/s": b32 / ms / 1e6, "frac_of_peak": b32 / ms / 1e6 / HBM_PEAK_GBS}

    cpu = None
    if not args.skip_cpu:
        from oracle import oracle as O
        t0 = time.perf_counter()
        if args.pressure_solver == 2:
            cpu = cpu_baseline(O, g64, st64, sz, 1, jacobi_pressure_iters(g64, args))
        else:
            cpu = cpu_baseline(O, g64, st64, sz, args.pressure_solver)
        log(f"cpu baseline in {time.perf_counter() - t0:.1f}s")
        if not args.skip_cpu_steps:
            t0 = time.perf_counter()
            cpu["whole_steps"] = {"config1_cylinder_L10": cpu_whole_steps(P, O, "config1", 5),
                                  "config2_cylinder_L30": cpu_whole_steps(P, O, "config2", 2)}
            log(f"cpu whole steps in {time.perf_counter() - t0:.1f}s: {cpu['whole_steps']}")

    value = res[primary]["steps_per_s"]
    out = {"metric": METRIC, "value": value, "unit": "PISO steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res[primary]["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if primary == "fp64" else "f32 inner / f64 refinement",
           "data": "synthetic",
           "config": {"workload": workload, "mesh": mesh_desc, "corrections": args.corrections,
                      "tets": sz["nt"], "velocity_dofs": 3 * sz["nu"],
                      "pressure_dofs": sz["nv"], "nnz_momentum": sz["nnzA"],
                      "nnz_pressure": sz["nnzP"], "density": args.rho, "viscosity": args.mu,
                      "inlet_peak": U, "reynolds": None if cyl else args.re, "dt": args.dt,
                      "pressure_tol": args.pressure_tol, "momentum_tol": args.momentum_tol,
                      "stop_rule": "jacobi_scaled" if args.stop_rule else "reference",
                      "pressure_solver": PRESSURE_SOLVERS[args.pressure_solver],
                      "fp64_momentum_solvers": {k: v["momentum_solver"] for k, v in res.items()
                                                if k in dict(FP64)},
                      "headline": primary,
                      "later_correctors": ("re-solve" if args.resolve_correctors else
                                           "reuse the first corrector's p' (identical inputs, "
                                           "same bits)") if args.corrections > 1 else None,
                      "velocity_error_bound": VEL_ERR_MAX, "parallelism": "single GPU"},
           "dofs_per_s": value * (3 * sz["nu"] + sz["nv"]),
           "pressure_amg": ({k: v for k, v in g64.amgInfo().items() if k != "builds"}
                            if args.pressure_solver == 2 else None),
           "host_readback_ms_per_step": host_copy_ms,
           "host_readback_pageable_ms_per_step": host_copy_pageable_ms,
           "steps_per_s_with_host_readback": 1e3 / (res[primary]["ms_per_step"] + host_copy_ms),
           "speedup_vs_fp64": (res[SAME_SOLVER[primary]]["ms_per_step"] /
                               res[primary]["ms_per_step"])
           if SAME_SOLVER.get(primary) in res else (1.0 if primary == "fp64" else None),
           "speedup_vs_fp64_solver": SAME_SOLVER.get(primary, "fp64"),
           "speedup_vs_fp64_bicgstab": res["fp64"]["ms_per_step"] / res[primary]["ms_per_step"],
           **res, "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
