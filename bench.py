"""Benchmark of the D3Q27 cumulant hot path on one MI355X.

Workload (the plain form on the GPU): uniform 256^3 box, FP32 populations, equilibrium velocity
inlet ux = 0.05 at x = 0, zero-gradient outlet, free-slip sides, sphere of diameter 32 cells centred at
(96, 128, 128) with halfway bounce-back, Re = u D / nu = 200 (nu = 0.008, tau = 0.524).  WALE and
inlet noise off (the plain parity form); cumulant defaults of the reference otherwise.  A step is one
fused stream-collide launch over the whole box; inputs are resident in HBM.

The same box with FP16 shifted storage (FP32 arithmetic) is timed as the precision form and reported
beside it.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 256
D = 32
U_IN = 0.05
RE = 200.0


def workload_grid():
    from openaerolab_alpha_01_amd.geometry import sphere_mask, uniform_grid
    g = uniform_grid((N, N, N))
    g.solid = sphere_mask(g, (96.0, 128.0, 128.0), D / 2)
    return g


def wall_qmap(grid):
    """wall distances of the sphere's links for interpolated bounce-back (--bouzidi)"""
    from openaerolab_alpha_01_amd.geometry import sphere_qmap
    return sphere_qmap(grid, (96.0, 128.0, 128.0), D / 2)


def tau_of():
    nu = U_IN * D / RE
    return 0.5 + 3.0 * nu


def bytes_per_lup(storage: str, wale: bool = False) -> int:
    """algorithmic HBM bytes per cell update: 27 populations read and written, plus with WALE the
    previous step's velocity read and this step's written (3 values of the working precision each)"""
    pops = 2 * 27 * {"f32": 4, "f64": 8, "f16s": 2, "bf16s": 2}[storage]
    return pops + (2 * 3 * (8 if storage == "f64" else 4) if wale else 0)


def gpu_leg(grid, storage: str, stream: str, steps: int, warmup: int, autotune: bool = True, fixed=None,
            qmap=None, wale: bool = False):
    """wale: the reference's production collision (WALE c_wale 0.2 with inlet noise 0.01,
    src/io/config_loader.jl:82-88) instead of the plain cumulant"""
    from openaerolab_alpha_01_amd import solver
    params = solver.StepParams(c_wale=0.2 if wale else 0.0, inlet_turbulence=0.01 if wale else 0.0)
    plan = solver.uniform_plan(grid, tau_of(), storage=storage, stream=stream, params=params, u_inlet=U_IN)
    if qmap is not None:
        plan.set_bouzidi(qmap)
    plan.init_equilibrium(None, None, 1.0, (U_IN, 0.0, 0.0))
    info = plan.level_info()
    tuning = None
    if autotune:
        best, table = plan.autotune(6)
        tuning = dict(best, ms_per_step={f"pipe{pp}_overlap{ov}_face{fk}": round(v, 4)
                                         for (pp, ov, fk), v in table.items()})
    else:
        if fixed:
            plan.set_tuning(**fixed)
        tuning = plan.tuning
    if warmup:
        plan.time_steps(warmup)
    plan.sync()
    t0 = time.perf_counter()
    ms_total, ms_kernel = plan.time_steps(steps)
    plan.sync()
    wall = time.perf_counter() - t0
    plan.level_step(0, steps + warmup, force=True)
    F = plan.force()[0]
    plan.guard(0.2, 0.5)
    plan.close()
    return ms_total, ms_kernel, wall, F, tuning, info


def copy_bandwidth_gbs() -> float:
    """liboal's streaming 16-byte copy of 4 GiB (read + write counted): the measured HBM ceiling."""
    import ctypes as C
    from openaerolab_alpha_01_amd import _lib
    lib = _lib.load()
    v = C.c_double()
    _lib.check(lib.oal_copy_bandwidth(0, 4 << 30, 10, C.byref(v)), "oal_copy_bandwidth")
    return v.value


def measured_traffic(kernel: str, summary: str = "step_f32_256_hbm_traffic.md"):
    """HBM bytes per launch of `kernel` (or per step of the launches joined by ' + ') from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None when the summary covers another
    kernel or launch shape."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    text = open(path).read()
    for k, b in re.findall(r"^\s*kernel = (.+)\n\s*hbm_bytes_per_launch = (\d+)$", text, re.M):
        if k.strip() == kernel:
            return int(b)
    return None


def cpu_baseline(grid, budget_s: float = 20.0):
    """The CPU oracle (this project's C++ restatement of the reference arithmetic, OpenMP) on the same
    workload, FP32: as many whole 256^3 steps as fit in the budget (at least one)."""
    from oracle import oracle as O
    G = O.Grid(grid)
    n = grid.n_cells
    rho = np.ones(n, dtype=np.float32)
    u = np.zeros((3, n), dtype=np.float32)
    u[0] = U_IN
    f = O.init_eq("f32", G, rho, u)
    g = f.copy()
    del rho, u
    p = O.make_params(tau_of(), u_inlet=U_IN, arith=O.LITERAL)
    done, t_used = 0, 0.0
    while done == 0 or (t_used < budget_s and done < 20):
        t0 = time.perf_counter()
        O.step("f32", G, p, f, out=g)
        t_used += time.perf_counter() - t0
        f, g = g, f
        done += 1
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    return n * done / t_used / 1e6, threads, done, t_used


def validation(steps: int, seed: int, chunk: int = 1000, bouzidi: bool = False) -> dict:
    """Accuracy of the storage forms on the same workload against an FP64 run of the same kernels (the
    FP64 path equals the FP64 oracle bit for bit in the parity tests).

    C2: FP32 populations from the uniform start.  C3: FP16 and BF16 shifted storage from a start whose
    velocity carries a seeded perturbation, uniform in [-1e-3, 1e-3] per component (numpy PCG64,
    `seed`), with the FP64 run started from the same field.  After `steps` steps (the last one
    collecting the momentum exchange): Cd = F_x / (1/2 rho0 u_in^2 pi R^2) and
    L2(u) = |u - u_64| / |u_64| over the fluid cells.  Beside the momentum exchange, the reference
    application's own drag estimate: the surface-stress map of the sphere's triangulated surface
    (5120 triangles of radius R; pressure and wall-law shear, src/forces/surface.jl, C_SMAGO 0.2) on the
    same final state, cd_surface.  bouzidi: the sphere's links take interpolated bounce-back (every
    storage form, the FP64 run included)."""
    from openaerolab_alpha_01_amd import solver
    from openaerolab_alpha_01_amd.geometry import icosphere
    grid = workload_grid()
    mesh = icosphere((96.0, 128.0, 128.0), D / 2, 4)
    qmap = wall_qmap(grid) if bouzidi else None
    fluid = grid.solid == 0
    n = grid.n_cells
    area = np.pi * (D / 2) ** 2
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    out = {"steps": steps, "seed": seed, "cd_reference_area": round(area, 4),
           "wall": "interpolated bounce-back (BFL)" if bouzidi else "halfway bounce-back"}
    for case, sd, forms in (("C2", None, ("f32",)), ("C3", seed, ("f16s", "bf16s"))):
        u0 = np.zeros((3, n))
        u0[0] = U_IN
        if sd is not None:
            u0 += np.random.default_rng(sd).uniform(-1e-3, 1e-3, size=(3, n))
        rho0 = np.ones(n)
        ref = None
        res = {}
        for st in ("f64",) + forms:
            t0 = time.perf_counter()
            plan = solver.uniform_plan(grid, tau_of(), storage=st, params=params, u_inlet=U_IN)
            if qmap is not None:
                plan.set_bouzidi(qmap)
            plan.init_equilibrium(rho0, u0)
            t = 0
            while t < steps - 1:
                k = min(chunk, steps - 1 - t)
                plan.step(k, t)
                plan.sync()
                t += k
                print(f"[validate] {case} {st} step {t}/{steps} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
                      flush=True)
            plan.level_step(0, t, force=True)
            F = plan.force()[0]
            plan.set_surface(*mesh)
            ts = time.perf_counter()
            surf = plan.surface_forces(tau=tau_of(), c_wale=0.2, rho_ref=1.0, u_ref=U_IN, area_ref=area,
                                       chord_ref=float(D), moment_center=(96.0, 128.0, 128.0))
            surf_ms = (time.perf_counter() - ts) * 1e3
            _, u = plan.macro()
            plan.guard(0.2, 0.5)
            plan.close()
            uf = u[:, fluid]
            del u
            cd = float(F[0] / (0.5 * U_IN ** 2 * area))
            entry = {"cd": round(cd, 6), "cl_y": round(float(F[1] / (0.5 * U_IN ** 2 * area)), 7),
                     "cl_z": round(float(F[2] / (0.5 * U_IN ** 2 * area)), 7),
                     "cd_surface": round(surf["Cd"], 6),
                     "cd_surface_pressure": round(float(surf["F_pressure"][0] / (0.5 * U_IN ** 2 * area)), 6),
                     "surface_mapped": surf["n_mapped"], "surface_ms": round(surf_ms, 3)}
            if ref is None:
                ref = (cd, uf, surf["Cd"])
            else:
                entry["cd_rel_err_vs_f64"] = float((cd - ref[0]) / ref[0])
                entry["l2_u_vs_f64"] = float(np.linalg.norm(uf - ref[1]) / np.linalg.norm(ref[1]))
                entry["cd_surface_rel_err_vs_f64"] = float((surf["Cd"] - ref[2]) / ref[2])
            res[st] = entry
        out[case] = res
    return out


# ---- Taylor-Green vortex, 64^3 periodic, FP64 (the reference's CPU-runnable case) ---------------------
TGV_N = 64
TGV_U0 = 0.05
TGV_RE = 1600.0


def tgv_fields(grid):
    """u = u0 sin x cos y cos z, v = -u0 cos x sin y cos z, w = 0 at x = 2 pi (i - 1/2) / 64 (cell
    centres), rho = 1 + 3 p' with p' = (u0^2 / 16)(cos 2x + cos 2y)(cos 2z + 2); nu = u0 L / Re with
    L = 64 (nu = 0.002, tau = 0.506)."""
    x = 2 * np.pi * (grid.cell_coords().astype(np.float64) + 0.5) / TGV_N
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    u = np.stack([TGV_U0 * np.sin(X) * np.cos(Y) * np.cos(Z), -TGV_U0 * np.cos(X) * np.sin(Y) * np.cos(Z),
                  np.zeros_like(X)])
    rho = 1.0 + 3.0 * (TGV_U0 ** 2 / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    return rho, u


def tgv_tau():
    return 0.5 + 3.0 * TGV_U0 * TGV_N / TGV_RE


def tgv_main(args) -> int:
    """C1: FP64 (and FP32) on the GPU, kinetic energy every 100 steps and MLUPS; the CPU oracle (FP64,
    literal form) on the same case, timed for a bounded number of steps, with its kinetic energy at
    those steps."""
    from openaerolab_alpha_01_amd import solver
    from openaerolab_alpha_01_amd.geometry import uniform_grid
    grid = uniform_grid((TGV_N,) * 3, periodic=(True, True, True))
    rho, u = tgv_fields(grid)
    n = grid.n_cells
    steps = args.tgv_steps
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    out = {"workload": f"Taylor-Green vortex, {TGV_N}^3 periodic, u0={TGV_U0}, Re={TGV_RE:g} on L={TGV_N} "
                       f"(nu={TGV_U0 * TGV_N / TGV_RE:g}, tau={tgv_tau():g}), equilibrium init, {steps} steps",
           "every": 100}
    for st in ("f64", "f32"):
        plan = solver.uniform_plan(grid, tgv_tau(), storage=st, params=params)
        plan.init_equilibrium(rho, u)
        ke, ms = [], 0.0
        for k in range(0, steps + 1, 100):
            r, v = plan.macro()
            ke.append(float(0.5 * np.sum(r * (v * v).sum(0))))
            if k < steps:
                tot, _ = plan.time_steps(min(100, steps - k))
                ms += tot
        plan.close()
        out[st] = {"ke": ke, "ms_total": round(ms, 3), "mlups": round(n * steps / (ms * 1e-3) / 1e6, 1)}
    out["f32_ke_rel_err_vs_f64"] = float(np.max(np.abs(np.array(out["f32"]["ke"]) / np.array(out["f64"]["ke"]) - 1)))
    # viscous decay rate of the initial mode, 2 nu |k|^2 with |k|^2 = 3 (2 pi / 64)^2, against the first 100 steps
    nu = (np.float32(tgv_tau()) - 0.5) / 3.0
    k2 = 3 * (2 * np.pi / TGV_N) ** 2
    ke = out["f64"]["ke"]
    # steps 0-100 include the start-up transient of the equilibrium initialisation
    out["decay_rate"] = {"steps_0_100": float(-np.log(ke[1] / ke[0]) / 100),
                         "steps_100_200": float(-np.log(ke[2] / ke[1]) / 100), "viscous_initial": float(2 * nu * k2)}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        G = O.Grid(grid)
        f = O.init_eq("f64", G, rho, u)
        g = f.copy()
        p = O.make_params(tgv_tau(), arith=O.LITERAL)
        cpu_steps = min(steps, args.tgv_cpu_steps)
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            O.step("f64", G, p, f, out=g)
            f, g = g, f
        secs = time.perf_counter() - t0
        _, r, v = O.moments("f64", G, f)
        ke_cpu = float(0.5 * np.sum(r * (v * v).sum(0)))
        threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        ke_gpu = out["f64"]["ke"][cpu_steps // 100] if cpu_steps % 100 == 0 else None
        out["cpu_baseline"] = {"value": round(n * cpu_steps / secs / 1e6, 2), "unit": "MLUPS", "cores": threads,
                               "kind": "port", "sample": f"{cpu_steps} FP64 steps of the same case in {secs:.1f} s "
                                                         "(CPU oracle, literal form)",
                               "ke": ke_cpu,
                               "ke_rel_diff_vs_gpu_f64": None if ke_gpu is None else float(ke_cpu / ke_gpu - 1)}
    print(json.dumps({"tgv": out}), flush=True)
    return 0


# ---- many-instance form: 48 independent boxes with ellipsoids, one launch per step ------------------
MANY_COUNT = 48
MANY_RE = 20.0       # Reynolds number of every box, u_in D / nu with D the ellipsoid's largest diameter (the
                     # configuration leaves nu open; a fixed tau of 0.52 to 0.60 let the faster, larger boxes'
                     # wakes reach the zero-gradient outlet unsteady, and some diverged within 5000 steps)


MANY_SPONGE = 0.15   # outlet sponge thickness in units of the reference length (advanced.domain.sponge_thickness)
MANY_SPONGE_SIGMA = 0.25   # its tanh transition width (sponge_tanh_sigma)


def many_sponge(inst) -> np.ndarray:
    """the outlet sponge of apply_sponge! (src/geometry/domain_generation.jl:296-349) for one box:
    thickness T = D max(sponge_thickness, 0.08) with D the ellipsoid's largest diameter, the tanh profile
    of tanh_sponge_profile (:265-275) in the distance from the outlet, outlet strength 1"""
    g = inst["grid"]
    nx = float(inst["n"][0])
    T = min(nx, 2.0 * float(np.max(inst["semi"])) * max(MANY_SPONGE, 0.08))
    xx = nx - (g.cell_coords()[:, 0] + 0.5)            # distance from the outlet
    s = np.where(xx <= 0.0, 1.0, np.where(xx >= T, 0.0,
                 0.5 * (1.0 - np.tanh((xx - 0.5 * T) / (MANY_SPONGE_SIGMA * T)))))
    s[~g.inside()] = 0.0
    return s.astype(np.float32)


def many_tau(inst) -> float:
    """relaxation time of a box at Re = MANY_RE: tau = 1/2 + 3 u_in D / Re"""
    return 0.5 + 3.0 * inst["u_in"] * 2.0 * float(np.max(inst["semi"])) / MANY_RE


def many_instances(seed: int):
    """48 boxes: edges uniform in [48, 160] cells per axis, an ellipsoid with semi-axes uniform in [6, 20]
    cells and a uniformly random orientation (unit quaternion), inlet velocity uniform in [0.02, 0.08];
    numpy PCG64 with `seed`.  The ellipsoid's centre sits at a third of the box length (at least its
    largest semi-axis + 4 cells from the inlet), mid-height and mid-width."""
    from openaerolab_alpha_01_amd.geometry import ellipsoid_mask, random_rotation, uniform_grid
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(MANY_COUNT):
        n = tuple(int(v) for v in rng.integers(48, 161, 3))
        semi = rng.uniform(6.0, 20.0, 3)
        R = random_rotation(rng)
        u_in = float(rng.uniform(0.02, 0.08))
        c = (max(n[0] / 3.0, semi.max() + 4.0), n[1] / 2.0, n[2] / 2.0)
        g = uniform_grid(n)
        g.solid = ellipsoid_mask(g, c, semi, R)
        out.append(dict(grid=g, n=n, semi=semi, R=R, u_in=u_in, center=c))
    return out


def many_plan(inst, storage):
    from openaerolab_alpha_01_amd import solver
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    plan = solver.batched_plan([i["grid"] for i in inst], [many_tau(i) for i in inst], [i["u_in"] for i in inst],
                               storage=storage, params=params)
    rho = np.concatenate([np.ones(i["grid"].n_cells) for i in inst])
    u = np.concatenate([np.stack([np.where(i["grid"].solid != 0, 0.0, i["u_in"]), np.zeros(i["grid"].n_cells),
                                  np.zeros(i["grid"].n_cells)]) for i in inst], axis=1)
    plan.set_sponge(np.concatenate([many_sponge(i) for i in inst]))
    plan.init_equilibrium(rho, u)
    return plan


def many_main(args) -> int:
    """C5: aggregate MLUPS of the 48 boxes in FP16 shifted storage (FP32 compute); with --validate, the
    per-box Cd after --many-steps steps against the FP64 run of the same kernels."""
    from openaerolab_alpha_01_amd.geometry import ellipsoid_frontal_area
    inst = many_instances(args.seed)
    cells = sum(int(np.prod(i["n"])) for i in inst)
    stored = sum(i["grid"].n_cells for i in inst)
    out = {"workload": f"{MANY_COUNT} independent inlet/outlet/free-slip boxes, edges U[48,160], ellipsoids with "
                       f"semi-axes U[6,20] at random orientation, u_in U[0.02,0.08], Re={MANY_RE:g} per box "
                       f"(tau = 1/2 + 3 u_in D / Re, D the largest diameter), outlet sponge of the reference's form "
                       f"(thickness {MANY_SPONGE:g} D, tanh sigma {MANY_SPONGE_SIGMA:g}), seed {args.seed}; "
                       "all boxes in one block list, one launch per kernel kind per step",
           "cells": cells, "stored_cells": stored,
           "tau_range": [round(min(many_tau(i) for i in inst), 4), round(max(many_tau(i) for i in inst), 4)]}
    if args.validate:
        res = {}
        for st in ("f64", "f16s"):
            t0 = time.perf_counter()
            plan = many_plan(inst, st)
            t = 0
            while t < args.many_steps - 1:
                k = min(1000, args.many_steps - 1 - t)
                plan.step(k, t)
                plan.sync()
                t += k
                print(f"[validate-many] {st} step {t}/{args.many_steps} ({time.perf_counter() - t0:.1f} s)",
                      file=sys.stderr, flush=True)
            plan.level_step(0, t, force=True)
            F = plan.force()
            mass = plan.mass()
            plan.close()
            res[st] = [float(F[k][0] / (0.5 * i["u_in"] ** 2 * ellipsoid_frontal_area(i["semi"], i["R"])))
                       for k, i in enumerate(inst)]
            res[st + "_finite"] = [bool(np.isfinite(mass[k]).all() and np.isfinite(F[k]).all())
                                   for k in range(len(inst))]
        ok = [a and b for a, b in zip(res["f64_finite"], res["f16s_finite"])]
        err = [(a - b) / b for a, b, o in zip(res["f16s"], res["f64"], ok) if o]
        # boxes that did not stay bounded, with what sets them apart: the gap between the body's far
        # end and the zero-gradient outlet, the relaxation time and the blockage
        out["unstable_boxes"] = [{"box": k, "f64_bounded": res["f64_finite"][k], "f16s_bounded": res["f16s_finite"][k],
                                  "tau": round(many_tau(i), 4), "u_in": round(i["u_in"], 4),
                                  "outlet_gap": round(i["n"][0] - i["center"][0] - float(np.max(i["semi"])), 1),
                                  "blockage": round(2 * float(np.max(i["semi"])) / min(i["n"][1], i["n"][2]), 3)}
                                 for k, (i, o) in enumerate(zip(inst, ok)) if not o]
        out["validation"] = {"steps": args.many_steps, "cd_f64": [round(v, 5) for v in res["f64"]],
                             "cd_f16s": [round(v, 5) for v in res["f16s"]],
                             "stable_boxes": int(sum(ok)),
                             "cd_rel_err": [float(e) for e in err],
                             "max_abs_cd_rel_err": float(np.max(np.abs(err))) if err else None,
                             "median_abs_cd_rel_err": float(np.median(np.abs(err))) if err else None}
    else:
        legs = []
        for st in ("f16s", "f32"):
            plan = many_plan(inst, st)
            tun = plan.tuning
            if not args.no_autotune:
                tun, _ = plan.autotune(6)
            plan.time_steps(args.warmup)
            ms, ker = plan.time_steps(args.steps)
            plan.close()
            legs.append({"storage": st, "ms_per_step": round(ms / args.steps, 4),
                         "mlups": round(cells * args.steps / (ms * 1e-3) / 1e6, 1),
                         "stored_mlups": round(stored * args.steps / (ms * 1e-3) / 1e6, 1), "tuning": tun})
        out["legs"] = legs
    print(json.dumps({"many": out}), flush=True)
    return 0


# ---- block-AMR form: octree of 16^3 blocks around a sphere -------------------------------------------
AMR_COARSE = (64, 32, 32)          # level-0 extent; finest-equivalent 512 x 256 x 256
AMR_LEVELS = 4                     # coarse + 3 refinement levels
AMR_CENTER = (192.0, 128.0, 128.0)  # finest cells
AMR_D = 48                         # sphere diameter, finest cells
AMR_THRESHOLDS = (64, 24, 8)       # distance to the surface (finest cells) refining into level 1, 2, 3


def amr_grids():
    from openaerolab_alpha_01_amd.geometry import sphere_octree
    return sphere_octree(AMR_COARSE, AMR_LEVELS, AMR_CENTER, AMR_D / 2, AMR_THRESHOLDS)


def amr_tau_finest():
    return 0.5 + 3.0 * U_IN * AMR_D / RE


def stepped_masks(grids):
    """per level, cell mask of the caller's blocks that are stepped (blocks whose children exist are not)"""
    out = []
    for l, g in enumerate(grids):
        blk = np.ones(g.n_blocks, dtype=bool)
        if l + 1 < len(grids):
            fine = {tuple(v) for v in (np.asarray(grids[l + 1].coords) - 1).tolist()}
            c0 = (np.asarray(g.coords, dtype=np.int64) - 1) * 2
            blk = np.array([tuple(c) not in fine for c in c0.tolist()])
        out.append(np.repeat(blk, 16 ** 3))
    return out


def amr_plan(grids, storage, u0_levels=None):
    from openaerolab_alpha_01_amd import solver
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    plan = solver.octree_plan(grids, amr_tau_finest(), storage=storage, params=params, u_inlet=U_IN)
    for l, g in enumerate(grids):
        if u0_levels is None:
            plan.init_equilibrium(None, None, 1.0, (U_IN, 0.0, 0.0), level=l)
        else:
            plan.init_equilibrium(np.ones(g.n_cells), u0_levels[l], level=l)
    return plan


def amr_leg(storage: str, steps: int, warmup: int, autotune: bool, graph: int):
    grids = amr_grids()
    plan = amr_plan(grids, storage)
    L = len(grids)
    tuning = None
    if autotune:
        best, _ = plan.autotune(6, level=L - 1)      # the finest level carries most of the work
        tuning = best
    plan.set_tuning(graph=graph)
    tuning = plan.tuning
    if warmup:
        plan.time_steps(warmup, level=None)
    plan.sync()
    t0 = time.perf_counter()
    ms_total, _ = plan.time_steps(steps, level=None)
    wall = time.perf_counter() - t0
    plan.step(1, steps + warmup, force=True)
    F = plan.force(level=L - 1)[0]
    bad = sum(plan.guard(0.2, 0.5, level=l) for l in range(L))
    infos = [plan.level_info(l) for l in range(L)]
    plan.close()
    masks = stepped_masks(grids)
    updates = sum(int(m.sum()) * 2 ** l for l, m in enumerate(masks))
    fe = int(np.prod(np.asarray(AMR_COARSE) << (L - 1))) * 2 ** (L - 1)
    return {"storage": storage, "graph": graph, "ms_per_coarse_step": round(ms_total / steps, 4),
            "mlups": round(updates * steps / (ms_total * 1e-3) / 1e6, 1),
            "finest_equivalent_mlups": round(fe * steps / (ms_total * 1e-3) / 1e6, 1),
            "cell_updates_per_coarse_step": updates, "wall_s": round(wall, 3), "tuning": tuning,
            "levels": [{k: i[k] for k in ("n_blocks", "n_covered_blocks", "n_ghost_blocks")} for i in infos],
            "cd_transient": round(float(F[0] / (0.5 * U_IN ** 2 * np.pi * (AMR_D / 2) ** 2)), 5), "guard_bad": bad}


def amr_validation(steps: int, seed: int, chunk: int = 250) -> dict:
    """Cd and L2(u) of the refined run in FP32 / FP16 / BF16 storage against the FP64 run of the same
    kernels, from the same start: u = (u_in, 0, 0) plus a seeded uniform [-1e-3, 1e-3] perturbation per
    component (numpy PCG64, `seed`, one stream per level).  steps are coarse steps (8 finest steps each);
    Cd from the finest level's momentum exchange, L2 over the stepped cells of every level."""
    grids = amr_grids()
    masks = stepped_masks(grids)
    area = np.pi * (AMR_D / 2) ** 2
    u0 = []
    for l, g in enumerate(grids):
        u = np.zeros((3, g.n_cells))
        u[0] = U_IN
        u += np.random.default_rng(seed + l).uniform(-1e-3, 1e-3, size=(3, g.n_cells))
        if g.solid is not None:
            u[:, np.asarray(g.solid) != 0] = 0.0
        u0.append(u)
    fluid = [m & (np.asarray(g.solid) == 0) for m, g in zip(masks, grids)]
    out = {"coarse_steps": steps, "finest_steps": steps * 2 ** (len(grids) - 1), "seed": seed,
           "cd_reference_area": round(area, 4)}
    ref = None
    for st in ("f64", "f32", "f16s", "bf16s"):
        t0 = time.perf_counter()
        plan = amr_plan(grids, st, u0)
        t = 0
        while t < steps - 1:
            k = min(chunk, steps - 1 - t)
            plan.step(k, t)
            plan.sync()
            t += k
            print(f"[validate-amr] {st} coarse step {t}/{steps} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
                  flush=True)
        plan.step(1, t, force=True)
        F = plan.force(level=len(grids) - 1)[0]
        us = np.concatenate([plan.macro(level=l)[1][:, fluid[l]] for l in range(len(grids))], axis=1)
        bad = sum(plan.guard(0.2, 0.5, level=l) for l in range(len(grids)))
        plan.close()
        cd = float(F[0] / (0.5 * U_IN ** 2 * area))
        entry = {"cd": round(cd, 6), "guard_bad": bad}
        if ref is None:
            ref = (cd, us)
        else:
            entry["cd_rel_err_vs_f64"] = float((cd - ref[0]) / ref[0])
            entry["l2_u_vs_f64"] = float(np.linalg.norm(us - ref[1]) / np.linalg.norm(ref[1]))
        out[st] = entry
    return out


def amr_main(args) -> int:
    """The block-AMR form: one JSON line (finest-equivalent MLUPS of FP16 storage, graph replay)."""
    if args.validate:
        print(json.dumps({"validation_amr": amr_validation(args.val_steps, args.seed)}), flush=True)
        return 0
    legs = []
    for item in args.amr_legs.split(","):
        st, gr = item.split(":")
        legs.append(amr_leg(st, args.steps, args.warmup, not args.no_autotune, int(gr)))
        print(f"[amr] leg {item}: {legs[-1]['ms_per_coarse_step']} ms per coarse step", file=sys.stderr, flush=True)
    head = legs[0]
    out = {"metric": "finest-equivalent MLUPS (D3Q27 cumulant block-AMR, FP16 shifted storage, FP32 compute)",
           "value": head["finest_equivalent_mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_coarse_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f16-storage/f32-compute", "data": "synthetic",
           "config": {"workload": f"block octree of 16^3 blocks, {AMR_LEVELS} levels (coarse "
                                  f"{'x'.join(map(str, AMR_COARSE))}, finest-equivalent "
                                  f"{'x'.join(str(v << (AMR_LEVELS - 1)) for v in AMR_COARSE)}), sphere D={AMR_D} finest "
                                  f"cells at {AMR_CENTER}, refinement thresholds {AMR_THRESHOLDS} finest cells, "
                                  f"acoustic sub-cycling, inlet ux={U_IN}, Re={RE:g}", "parallelism": "1 GPU"},
           "legs": legs}
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--stream", default="ab", choices=["ab", "aa"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-copy-bw", action="store_true", help="skip the copy-bandwidth reference measurement")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--pipe-planes", type=int, default=None, choices=[1, 2, 4, 8, 16],
                    help="fixed planes per workgroup of the bulk kernel (implies --no-autotune)")
    ap.add_argument("--overlap", type=int, default=None, choices=[0, 1],
                    help="fixed: boundary launches on a second stream (1) or in order (0) (implies --no-autotune)")
    ap.add_argument("--face-kernel", type=int, default=None, choices=[0, 1],
                    help="fixed: face cells by a separate kernel (1) or by the bulk kernel's rules (0) "
                         "(implies --no-autotune)")
    ap.add_argument("--validate", action="store_true",
                    help="accuracy run instead of the benchmark: Cd and L2(u) of FP32 / FP16 / BF16 vs FP64")
    ap.add_argument("--val-steps", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--bouzidi", action="store_true",
                    help="interpolated (BFL) bounce-back on the sphere's links instead of halfway bounce-back "
                         "(the default leg and --validate)")
    ap.add_argument("--amr", action="store_true",
                    help="the block-AMR form (4-level sphere octree) instead of the uniform box; with --validate "
                         "its accuracy run (--val-steps counts coarse steps)")
    ap.add_argument("--many", action="store_true",
                    help="the many-instance form (48 boxes with ellipsoids, FP16 storage) instead; with --validate "
                         "its per-box Cd against FP64 after --many-steps steps")
    ap.add_argument("--many-steps", type=int, default=5000)
    ap.add_argument("--tgv", action="store_true", help="the Taylor-Green vortex case (64^3 periodic, FP64) instead")
    ap.add_argument("--tgv-steps", type=int, default=1000)
    ap.add_argument("--tgv-cpu-steps", type=int, default=200, help="steps of the CPU oracle in --tgv")
    ap.add_argument("--amr-legs", default="f16s:1,f16s:0,f32:1",
                    help="block-AMR legs timed, storage:graph comma-separated; the first is the reported value")
    ap.add_argument("--no-aa-leg", action="store_true",
                    help="skip the in-place AA-pattern leg (FP32, one population buffer) of the default line")
    ap.add_argument("--no-wale-leg", action="store_true",
                    help="skip the WALE-on leg (FP32, c_wale 0.2, inlet noise 0.01) of the default line")
    ap.add_argument("--acc-steps", type=int, default=1500,
                    help="steps of the accuracy leg of the default line (Cd and L2(u) of FP32 / FP16 / BF16 against "
                         "FP64 runs of the same kernels); 0 skips it")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is a design, not an implementation)", file=sys.stderr)
        return 2
    fixed = {k: v for k, v in (("pipe_planes", args.pipe_planes), ("overlap", args.overlap),
                               ("face_kernel", args.face_kernel)) if v is not None}
    if fixed:
        args.no_autotune = True
    if args.amr:
        return amr_main(args)
    if args.tgv:
        return tgv_main(args)
    if args.many:
        return many_main(args)
    if args.validate:
        print(json.dumps({"validation": validation(args.val_steps, args.seed, bouzidi=args.bouzidi)}), flush=True)
        return 0
    if args.bouzidi:
        args.no_cpu_baseline = True   # the CPU baseline times the halfway form
    # the AA pattern steps in even/odd pairs: its counts are rounded up to even (and reported as run)
    warmup_asked = args.warmup
    if args.stream == "aa" and (args.steps % 2 or args.warmup % 2):
        args.steps += args.steps % 2
        args.warmup += args.warmup % 2

    grid = workload_grid()
    n_cells = N ** 3
    n_fluid = int(n_cells - grid.solid.sum())
    qmap = wall_qmap(grid) if args.bouzidi else None
    wall_rule = "interpolated (BFL) bounce-back" if args.bouzidi else "halfway bounce-back"
    ms_total, ms_kernel, wall, F, tun, info = gpu_leg(grid, "f32", args.stream, args.steps, args.warmup,
                                                      not args.no_autotune, fixed, qmap)
    mlups = n_cells * args.steps / (ms_total * 1e-3) / 1e6
    ms_step = ms_total / args.steps
    pp = tun["pipe_planes"]
    # dominant kernel: the bulk launch over the regular blocks (every cell fluid); the face cells are
    # stepped by it when it applies the face rules itself, else by the face-cell kernel
    face_in_bulk = args.stream == "ab" and pp > 1 and not tun["face_kernel"]
    bulk_cells = (info["n_regular"] + info["n_interior"]) * 16 ** 3 - (0 if face_in_bulk else info["n_face_cells"])
    achieved = bytes_per_lup("f32") * bulk_cells / (ms_kernel * 1e-3) / 1e9
    step_gbs = bytes_per_lup("f32") * n_fluid / (ms_step * 1e-3) / 1e9
    area = np.pi * (D / 2) ** 2
    cd = float(F[0] / (0.5 * U_IN ** 2 * area))

    # (the collision's unit-rate form: the reference's default rates omega_bulk = omega_3 = omega_4 = 1)
    if args.stream == "ab" and pp > 1 and tun.get("pair_kernel") and not tun["face_kernel"]:
        kernel = (f"k_step_pair32<{pp}, true, 1> + k_step_pair32<{pp}, true, 0>" if tun["pair_kernel"] == 2
                  else f"k_step_pair32<{pp}, true, 1>")
    elif args.stream == "ab" and pp > 1:
        kernel = f"k_step_pipe<StF32, {pp}, {2 if tun['face_kernel'] else 1}, true, false>"
    else:
        kernel = "k_step<StF32, regular>"
    traffic = measured_traffic(kernel)

    out = {
        "metric": "MLUPS (D3Q27 cumulant, FP32 populations, 256^3 sphere box)",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup_asked,
        "ms_per_step": round(ms_total / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "256^3 box, inlet ux=0.05, zero-gradient outlet, free-slip sides, sphere D=32 at "
                               f"(96,128,128), {wall_rule}, Re=200 (tau=0.524), cumulant, WALE off",
                   "cells": n_cells, "fluid_cells": n_fluid, "stream": args.stream, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel,
                     "kernel_ms": round(ms_kernel, 5), "cells_per_launch": bulk_cells,
                     "bytes_per_lup": bytes_per_lup("f32"),
                     "step_ms": round(ms_step, 5), "step_achieved_gbs": round(step_gbs, 1),
                     "step_frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "blocks": {k: info[k] for k in ("n_blocks", "n_regular", "n_interior", "n_boundary_regular", "n_boundary",
                                        "n_face_cells")},
        "cd_transient": round(cd, 5),
        "tuning": tun,
    }
    if not args.no_fp16:
        ms_t16, ms_k16, _, F16, tun16, _ = gpu_leg(grid, "f16s", args.stream, args.steps, args.warmup,
                                                   not args.no_autotune, fixed, qmap)
        m16 = n_cells * args.steps / (ms_t16 * 1e-3) / 1e6
        face16 = args.stream == "ab" and tun16["pipe_planes"] > 1 and not tun16["face_kernel"]
        bulk16 = (info["n_regular"] + info["n_interior"]) * 16 ** 3 - (0 if face16 else info["n_face_cells"])
        a16 = bytes_per_lup("f16s") * bulk16 / (ms_k16 * 1e-3) / 1e9
        s16 = bytes_per_lup("f16s") * n_fluid / (ms_t16 / args.steps * 1e-3) / 1e9
        pp16 = tun16["pipe_planes"]
        if args.stream == "ab" and pp16 > 1 and tun16.get("pair_kernel") and not tun16["face_kernel"]:
            k16 = (f"k_step_pair<StF16S, {pp16}, true, 1> + k_step_pair<StF16S, {pp16}, true, 0>"
                   if tun16["pair_kernel"] == 2 else f"k_step_pair<StF16S, {pp16}, true, 1>")
        elif args.stream == "ab" and pp16 > 1:
            k16 = f"k_step_pipe<StF16S, {pp16}, {2 if tun16['face_kernel'] else 1}, true, false>"
        else:
            k16 = "k_step<StF16S, regular>"
        traffic16 = measured_traffic(k16, "step_f16_256_hbm_traffic.md")
        out["fp16_shifted"] = {"value": round(m16, 1), "unit": "MLUPS", "ms_per_step": round(ms_t16 / args.steps, 4),
                               "speedup_vs_f32": round(m16 / mlups, 3), "tuning": tun16,
                               "roofline": {"achieved": round(a16, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                            "frac": round(a16 / HBM_PEAK_GBS, 4), "traffic": traffic16,
                                            "kernel": k16, "kernel_ms": round(ms_k16, 5),
                                            "cells_per_launch": bulk16,
                                            "bytes_per_lup": bytes_per_lup("f16s"),
                                            "step_achieved_gbs": round(s16, 1),
                                            "step_frac": round(s16 / HBM_PEAK_GBS, 4)}}
    if not args.no_copy_bw:
        bw = copy_bandwidth_gbs()
        out["roofline"]["measured_copy_gbs"] = round(bw, 1)
        out["roofline"]["frac_of_measured_copy"] = round(achieved / bw, 4)
    if not args.no_aa_leg and args.stream == "ab" and qmap is None:
        # the in-place AA pattern on the same workload: one population buffer instead of two (halfway
        # bounce-back only: interpolated bounce-back needs the AB pattern)
        aa_steps, aa_warm = args.steps + args.steps % 2, args.warmup + args.warmup % 2
        ms_ta, ms_ka, _, Fa, tuna, infoa = gpu_leg(grid, "f32", "aa", aa_steps, aa_warm, not args.no_autotune, fixed, qmap)
        ma = n_cells * aa_steps / (ms_ta * 1e-3) / 1e6
        sa = bytes_per_lup("f32") * n_fluid / (ms_ta / aa_steps * 1e-3) / 1e9
        out["aa_pattern"] = {"value": round(ma, 1), "unit": "MLUPS", "steps": aa_steps,
                             "ms_per_step": round(ms_ta / aa_steps, 4), "vs_ab": round(ma / mlups, 3),
                             "population_bytes": 27 * 4 * n_cells, "population_bytes_ab": 2 * 27 * 4 * n_cells,
                             "step_achieved_gbs": round(sa, 1), "step_frac": round(sa / HBM_PEAK_GBS, 4),
                             "tuning": {k: tuna[k] for k in ("pipe_planes", "overlap", "face_kernel")},
                             "blocks": {"pipelined_regular": infoa["n_regular"],
                                        "one_plane_boundary": infoa["n_boundary_regular"] + infoa["n_boundary"]},
                             "cd_transient": round(float(Fa[0] / (0.5 * U_IN ** 2 * area)), 5)}
    if not args.no_wale_leg:
        # the reference's production collision on the same workload: WALE (c_wale 0.2) and inlet noise (0.01)
        ms_tw, ms_kw, _, Fw, tunw, _ = gpu_leg(grid, "f32", args.stream, args.steps, args.warmup,
                                               not args.no_autotune, fixed, qmap, wale=True)
        mw = n_cells * args.steps / (ms_tw * 1e-3) / 1e6
        facew = args.stream == "ab" and tunw["pipe_planes"] > 1
        bulkw = (info["n_regular"] + info["n_interior"]) * 16 ** 3 - (0 if facew else info["n_face_cells"])
        bw = bytes_per_lup("f32", wale=True)
        aw = bw * bulkw / (ms_kw * 1e-3) / 1e9
        ppw = min(tunw["pipe_planes"], 8)
        kw = f"k_step_pipe<StF32, {ppw}, 1, true, true>" if facew else "k_step<StF32, regular, wale>"
        out["wale_on"] = {"value": round(mw, 1), "unit": "MLUPS", "ms_per_step": round(ms_tw / args.steps, 4),
                          "vs_wale_off": round(mw / mlups, 3), "c_wale": 0.2, "inlet_turbulence": 0.01,
                          "tuning": tunw, "cd_transient": round(float(Fw[0] / (0.5 * U_IN ** 2 * area)), 5),
                          "roofline": {"achieved": round(aw, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": round(aw / HBM_PEAK_GBS, 4), "bytes_per_lup": bw,
                                       "traffic": measured_traffic(kw, "step_wale_256_hbm_traffic.md"),
                                       "kernel": kw,
                                       "kernel_ms": round(ms_kw, 5), "cells_per_launch": bulkw}}
    if args.acc_steps > 0:
        # bounded accuracy leg: the storage forms against FP64 runs of the same kernels after acc_steps steps
        # (C2 FP32 from the uniform start, C3 FP16 / BF16 from the seeded perturbation; a transient state,
        # the steady-state record is --validate)
        acc = validation(args.acc_steps, args.seed, bouzidi=args.bouzidi)
        out["accuracy"] = {"steps": acc["steps"], "seed": acc["seed"], "wall": acc["wall"],
                           "cd_f64_C2": acc["C2"]["f64"]["cd"], "cd_f64_C3": acc["C3"]["f64"]["cd"],
                           "cd_rel_err_vs_f64": {st: round(acc[c][st]["cd_rel_err_vs_f64"], 7)
                                                 for c, st in (("C2", "f32"), ("C3", "f16s"), ("C3", "bf16s"))},
                           "l2_u_vs_f64": {st: float(f"{acc[c][st]['l2_u_vs_f64']:.3e}")
                                           for c, st in (("C2", "f32"), ("C3", "f16s"), ("C3", "bf16s"))},
                           # the reference application's drag convention on the same states (surface.jl)
                           "cd_surface_f64_C2": acc["C2"]["f64"]["cd_surface"],
                           "cd_surface_f64_C3": acc["C3"]["f64"]["cd_surface"],
                           "cd_surface_rel_err_vs_f64": {st: round(acc[c][st]["cd_surface_rel_err_vs_f64"], 7)
                                                         for c, st in (("C2", "f32"), ("C3", "f16s"), ("C3", "bf16s"))},
                           "surface_mapped_triangles": int(acc["C2"]["f64"]["surface_mapped"]),
                           "surface_map_ms": acc["C2"]["f32"]["surface_ms"]}
    if not args.no_cpu_baseline:
        v, cores, done, secs = cpu_baseline(grid)
        out["cpu_baseline"] = {"value": round(v, 2), "unit": "MLUPS", "cores": cores, "kind": "port",
                               "sample": f"{done} FP32 steps of the same 256^3 workload in {secs:.1f} s "
                                         "(CPU oracle: C++/OpenMP restatement of the reference arithmetic)"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
