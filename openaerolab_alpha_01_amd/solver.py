"""Host side of the MI355X D3Q27 cumulant hot path, mirroring the reference's solver entry points.

Reference interface (paths relative to the r-acad/OpenAeroLab_alpha_01 tree):
  get_collision_kernel      src/collision/physics_kernels.jl:50-74   -> get_collision_kernel
  perform_timestep_v2!      src/collision/timestep.jl:946-2000       -> perform_timestep_v2
  execute_timestep_batch!   src/solver/solver_control.jl:796-824     -> execute_timestep_batch
  init_eq!                  src/main.jl:546-572                      -> Plan.init_equilibrium
  BlockLevel                src/geometry/blocks.jl:16-200            -> geometry.BlockGrid + Domain

Every call goes through liboal.so (include/oal.h); there is no host or fallback arithmetic here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, load
from .geometry import BlockGrid


@dataclass
class StepParams:
    """Collision parameters with the reference defaults (src/io/config_loader.jl:82-88, 154-159, 223).

    The reference defaults turn WALE (c_wale 0.2) and inlet noise (0.01) on; the parity and
    benchmark cases of this project set both to 0 explicitly."""
    omega_bulk: float = 1.0
    omega_3: float = 1.0
    omega_4: float = 1.0
    lambda_: float = 1.0 / 6.0
    adaptive_omega4: int = 1
    limiter: int = _lib.OAL_LIMITER_FACTORED
    compressible: int = 0
    c_wale: float = 0.20
    nu_sgs_background: float = 0.0
    inlet_turbulence: float = 0.01
    transverse_freestream: int = 0          # advanced.boundary.transverse_freestream (config_loader.jl:2425)
    transverse_freestream_blend: float = 1.0

    def to_c(self) -> _lib.OalCollision:
        return _lib.OalCollision(self.omega_bulk, self.omega_3, self.omega_4, self.lambda_, self.adaptive_omega4,
                                 self.limiter, self.compressible, self.c_wale, self.nu_sgs_background,
                                 self.inlet_turbulence, int(self.transverse_freestream),
                                 self.transverse_freestream_blend)


@dataclass
class Domain:
    """One domain of a level: extent, boundary conditions and relaxation time."""
    extent: tuple[int, int, int]
    tau: float
    periodic: tuple[bool, bool, bool] = (False, False, False)
    u_inlet: float = 0.0
    body_force: tuple[float, float, float] = (0.0, 0.0, 0.0)
    amr_sgs_scale: float = 1.0
    ymin_wall_bc: int = 0
    zwall_bc: int = 0
    force_scale: float = 1.0

    def to_c(self) -> _lib.OalDomainDesc:
        d = _lib.OalDomainDesc()
        d.extent[:] = [int(v) for v in self.extent]
        d.periodic[:] = [1 if p else 0 for p in self.periodic]
        d.tau = self.tau
        d.u_inlet = self.u_inlet
        d.body_force[:] = list(self.body_force)
        d.amr_sgs_scale = self.amr_sgs_scale
        d.ymin_wall_bc = self.ymin_wall_bc
        d.zwall_bc = self.zwall_bc
        d.force_scale = self.force_scale
        return d


OPERATORS = {"cumulant"}


def get_collision_kernel(operator: str = "cumulant", lattice: str = "d3q27") -> str:
    """The operator registry of the reference (physics_kernels.jl:50-74), restricted to this path:
    :cumulant on D3Q27 is served by liboal; every other operator is outside it."""
    op = operator.lstrip(":")
    if op == "cumulant":
        if lattice != "d3q27":
            raise ValueError("[Collision] Cumulant operator is not compatible with D3Q39 lattice. "
                             "The Chimera transform requires 3x3x3 (D3Q27) structure.")
        return "cumulant"
    if op == "entropic":
        raise NotImplementedError("[Collision] collision_operator: entropic (D3Q39) is outside this library")
    raise ValueError(f"[Collision] Unknown or retired operator: :{op}. Valid options: :cumulant (D3Q27).")


@dataclass
class Level:
    grid: BlockGrid
    domains: list[Domain]
    block_domain: np.ndarray | None = None
    _keep: list = field(default_factory=list)


class Plan:
    """A solver run on the GPU: the device state of every level and the step entry points."""

    def __init__(self, levels: list[Level] | Level, storage: str = "f32", stream: str = "ab",
                 bb_mode: str = "halfway", params: StepParams | None = None, device: int = 0,
                 rounding: str = "stochastic"):
        self.lib = load()
        if isinstance(levels, Level):
            levels = [levels]
        self.levels = levels
        self.storage = storage
        self.stream_mode = stream
        self.params = params or StepParams()
        desc = _lib.OalPlanDesc()
        desc.abi_version = _lib.OAL_ABI_VERSION
        desc.storage = _lib.STORAGE_CODES[storage]
        desc.block_size = levels[0].grid.bs
        desc.stream_mode = {"ab": _lib.OAL_STREAM_AB, "aa": _lib.OAL_STREAM_AA}[stream]
        desc.bb_mode = {"reference": _lib.OAL_BB_REFERENCE, "halfway": _lib.OAL_BB_HALFWAY}[bb_mode]
        desc.n_levels = len(levels)
        descs = (_lib.OalLevelDesc * len(levels))()
        self._keep = []
        for i, lv in enumerate(levels):
            g = lv.grid
            mx = np.ascontiguousarray(g.coords[:, 0], dtype=np.int32)
            my = np.ascontiguousarray(g.coords[:, 1], dtype=np.int32)
            mz = np.ascontiguousarray(g.coords[:, 2], dtype=np.int32)
            nt = np.asfortranarray(g.neighbor, dtype=np.int32).ravel(order="F")   # column-major nb x 27
            doms = (_lib.OalDomainDesc * len(lv.domains))(*[d.to_c() for d in lv.domains])
            self._keep += [mx, my, mz, nt, doms]
            ld = descs[i]
            ld.n_blocks = g.n_blocks
            ld.map_x = mx.ctypes.data_as(C.POINTER(C.c_int32))
            ld.map_y = my.ctypes.data_as(C.POINTER(C.c_int32))
            ld.map_z = mz.ctypes.data_as(C.POINTER(C.c_int32))
            ld.neighbor_table = nt.ctypes.data_as(C.POINTER(C.c_int32))
            if lv.block_domain is not None:
                bd = np.ascontiguousarray(lv.block_domain, dtype=np.int32)
                self._keep.append(bd)
                ld.block_domain = bd.ctypes.data_as(C.POINTER(C.c_int32))
            ld.n_domains = len(lv.domains)
            ld.domains = C.cast(doms, C.POINTER(_lib.OalDomainDesc))
            if g.solid is not None:
                so = np.ascontiguousarray(g.solid, dtype=np.uint8)
                self._keep.append(so)
                ld.obstacle = so.ctypes.data_as(C.POINTER(C.c_uint8))
        desc.levels = C.cast(descs, C.POINTER(_lib.OalLevelDesc))
        desc.collision = self.params.to_c()
        desc.device = device
        desc.rounding = {"stochastic": _lib.OAL_ROUND_STOCHASTIC, "nearest": _lib.OAL_ROUND_NEAREST}[rounding]
        self._keep.append(descs)
        h = C.c_void_p()
        check(self.lib.oal_plan_create(C.byref(desc), C.byref(h)), "oal_plan_create")
        self.h = h

    # -- lifetime ----------------------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.oal_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------------------------------
    def n_cells(self, level: int = 0) -> int:
        n = C.c_int64()
        check(self.lib.oal_level_cells(self.h, level, C.byref(n)), "oal_level_cells")
        return n.value

    def upload_f(self, f: np.ndarray, level: int = 0) -> None:
        """f: [27, n_cells] (or flat) unshifted populations in the reference layout, float32 or float64."""
        f = np.ascontiguousarray(f)
        dt = _lib.OAL_F64 if f.dtype == np.float64 else _lib.OAL_F32
        f = f.astype(np.float64 if dt == _lib.OAL_F64 else np.float32, copy=False)
        if f.size != 27 * self.n_cells(level):
            raise ValueError("population array size does not match the level")
        check(self.lib.oal_upload_f(self.h, level, f.ctypes.data, dt), "oal_upload_f")

    def download_f(self, level: int = 0, dtype=np.float64) -> np.ndarray:
        out = np.empty((27, self.n_cells(level)), dtype=dtype)
        dt = _lib.OAL_F64 if out.dtype == np.float64 else _lib.OAL_F32
        check(self.lib.oal_download_f(self.h, level, out.ctypes.data, dt), "oal_download_f")
        return out

    def init_equilibrium(self, rho: np.ndarray | None = None, u: np.ndarray | None = None, rho0: float = 1.0,
                         u0=(0.0, 0.0, 0.0), level: int = 0) -> None:
        """init_eq! at per-cell (rho [n_cells], u [3, n_cells]) or uniform (rho0, u0)."""
        dp = C.POINTER(C.c_double)
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        v = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        u0a = np.ascontiguousarray(u0, dtype=np.float64)
        check(self.lib.oal_init_equilibrium(self.h, level, None if r is None else r.ctypes.data_as(dp),
                                            None if v is None else v.ctypes.data_as(dp), rho0,
                                            u0a.ctypes.data_as(dp)), "oal_init_equilibrium")

    # -- stepping ----------------------------------------------------------------------------------
    def level_step(self, level: int = 0, time_step: int = 0, force: bool = False) -> None:
        check(self.lib.oal_level_step(self.h, level, time_step, _lib.OAL_STEP_FORCE if force else 0, None),
              "oal_level_step")

    def step(self, n: int, first_time_step: int = 0, force: bool = False) -> None:
        check(self.lib.oal_step(self.h, n, first_time_step, _lib.OAL_STEP_FORCE if force else 0, None), "oal_step")

    def sync(self) -> None:
        check(self.lib.oal_sync(self.h), "oal_sync")

    def level_info(self, level: int = 0) -> dict:
        """Cell counts of a level and how its blocks split over the step kernels."""
        info = _lib.OalLevelInfo()
        check(self.lib.oal_level_info(self.h, level, C.byref(info)), "oal_level_info")
        return {name: getattr(info, name) for name, _ in _lib.OalLevelInfo._fields_ if name != "reserved"}

    def time_steps(self, n: int, level: int = 0, force: bool = False) -> tuple[float, float]:
        """Times n steps with HIP events: (ms for all n steps, mean ms of one step's bulk launches).
        level None: n whole (coarse) steps of every level through oal_step: (ms, mean ms per step)."""
        level = _lib.OAL_ALL_LEVELS if level is None else level
        tot, ker = C.c_double(), C.c_double()
        check(self.lib.oal_time_steps(self.h, level, n, _lib.OAL_STEP_FORCE if force else 0, C.byref(tot),
                                      C.byref(ker)), "oal_time_steps")
        return tot.value, ker.value

    # -- launch shape ------------------------------------------------------------------------------
    @property
    def tuning(self) -> dict:
        """Launch shape: pipe_planes (1, 2, 4, 8, 16), overlap (0 or 1), face_kernel (0 or 1), graph (0 or 1:
        step() replays a captured graph of one coarse step)."""
        t = _lib.OalTuning()
        check(self.lib.oal_get_tuning(self.h, C.byref(t)), "oal_get_tuning")
        return {"pipe_planes": t.pipe_planes, "overlap": t.overlap, "face_kernel": t.face_kernel, "graph": t.graph,
                "pair_kernel": t.pair_kernel}

    def set_tuning(self, pipe_planes: int | None = None, overlap: int | None = None,
                   face_kernel: int | None = None, graph: int | None = None, pair_kernel: int | None = None) -> None:
        cur = self.tuning
        t = _lib.OalTuning()
        t.pipe_planes = int(cur["pipe_planes"] if pipe_planes is None else pipe_planes)
        t.overlap = int(cur["overlap"] if overlap is None else overlap)
        t.face_kernel = int(cur["face_kernel"] if face_kernel is None else face_kernel)
        t.graph = int(cur["graph"] if graph is None else graph)
        t.pair_kernel = int(cur["pair_kernel"] if pair_kernel is None else pair_kernel)
        check(self.lib.oal_set_tuning(self.h, C.byref(t)), "oal_set_tuning")

    BFL_FALLBACKS = {"halfway_bb": _lib.OAL_BFL_FALLBACK_HALFWAY, "wall_eq": _lib.OAL_BFL_FALLBACK_WALL_EQ,
                     "bounce_back": _lib.OAL_BFL_FALLBACK_BOUNCE_BACK}

    def set_bouzidi(self, q_map, ff_fallback: str = "halfway_bb", q_min: float = 0.001, level: int = 0) -> None:
        """Interpolated bounce-back on the solid links of a level (apply_bouzidi_correction!,
        src/geometry/bouzidi_kernel.jl:382-432, in pull form).  q_map: [27, n_cells] wall distances
        (float16, or their uint16 bits; e.g. geometry.sphere_qmap), None to return to halfway
        bounce-back.  ff_fallback: "halfway_bb", "wall_eq" or "bounce_back" (advanced.boundary.ff_fallback,
        src/io/config_loader.jl:2519-2527)."""
        if ff_fallback not in self.BFL_FALLBACKS:
            raise ValueError(f"ff_fallback must be one of {sorted(self.BFL_FALLBACKS)}")
        ptr = None
        if q_map is not None:
            q = np.asarray(q_map)
            if q.dtype != np.uint16:
                q = q.astype(np.float16).view(np.uint16)
            n = self.n_cells(level)
            if q.shape != (27, n):
                raise ValueError(f"q_map must be [27, {n}]")
            q = np.ascontiguousarray(q)   # copied by the call
            ptr = q.ctypes.data
        check(self.lib.oal_set_bouzidi(self.h, level, ptr, self.BFL_FALLBACKS[ff_fallback], C.c_float(q_min)),
              "oal_set_bouzidi")

    def set_sponge(self, sponge, blend_distributions: bool = False, level: int = 0) -> None:
        """Sponge layer of a level (kernel_cumulant.jl:428-446): per-cell strengths [n_cells] in [0, 1] in the
        reference layout, or None to remove it; blend_distributions also blends the populations
        (SPONGE_BLEND_DISTRIBUTIONS)."""
        if sponge is None:
            check(self.lib.oal_set_sponge(self.h, level, None, 0), "oal_set_sponge")
            return
        sp = np.ascontiguousarray(sponge, dtype=np.float32)
        if sp.shape != (self.n_cells(level),):
            raise ValueError(f"sponge must be [{self.n_cells(level)}]")
        check(self.lib.oal_set_sponge(self.h, level, sp.ctypes.data, int(bool(blend_distributions))), "oal_set_sponge")

    def write_raw_state(self, path, level: int = 0, level_id: int | None = None, dx: float = 1.0) -> None:
        """The level's density and velocity as the reference's LBM1 raw state file (dump_raw_state,
        src/io/raw_state_dump.jl:21-47); level_id defaults to the reference's 1-based level + 1."""
        lid = level + 1 if level_id is None else int(level_id)
        check(self.lib.oal_write_raw_state(self.h, level, lid, float(dx), str(path).encode()), "oal_write_raw_state")

    def read_raw_state(self, path, level: int = 0) -> tuple[int, float]:
        """Start the level at the equilibrium of an LBM1 raw state file of the same blocks and obstacle;
        returns the file's (level_id, dx)."""
        lid = C.c_int32(0)
        dx = C.c_double(0.0)
        check(self.lib.oal_read_raw_state(self.h, level, str(path).encode(), C.byref(lid), C.byref(dx)),
              "oal_read_raw_state")
        return int(lid.value), float(dx.value)

    @property
    def pipe_planes(self) -> int:
        return self.tuning["pipe_planes"]

    @pipe_planes.setter
    def pipe_planes(self, v: int) -> None:
        self.set_tuning(pipe_planes=v)

    def autotune(self, n_steps: int = 10, level: int = 0) -> tuple[dict, dict]:
        """Time every launch shape, keep the fastest; the populations are restored afterwards.
        Returns the choice and {(pipe_planes, overlap, face_kernel): ms per step}."""
        best = _lib.OalTuning()
        ms = (C.c_double * _lib.OAL_TUNE_SHAPES)()
        check(self.lib.oal_autotune(self.h, level, n_steps, C.byref(best), ms), "oal_autotune")
        choice = {"pipe_planes": best.pipe_planes, "overlap": best.overlap, "face_kernel": best.face_kernel,
                  "graph": best.graph, "pair_kernel": best.pair_kernel}
        return choice, dict(zip(_lib.TUNE_KEYS, list(ms)))

    # -- results -----------------------------------------------------------------------------------
    def force(self, level: int = 0) -> np.ndarray:
        nd = len(self.levels[level].domains)
        out = np.empty(3 * nd, dtype=np.float64)
        check(self.lib.oal_force(self.h, level, out.ctypes.data_as(C.POINTER(C.c_double))), "oal_force")
        return out.reshape(nd, 3)

    def mass(self, level: int = 0) -> np.ndarray:
        nd = len(self.levels[level].domains)
        out = np.empty(4 * nd, dtype=np.float64)
        check(self.lib.oal_mass(self.h, level, out.ctypes.data_as(C.POINTER(C.c_double)), None), "oal_mass")
        return out.reshape(nd, 4)

    def macro(self, level: int = 0) -> tuple[np.ndarray, np.ndarray]:
        n = self.n_cells(level)
        rho = np.empty(n, dtype=np.float64)
        u = np.empty((3, n), dtype=np.float64)
        dp = C.POINTER(C.c_double)
        check(self.lib.oal_download_macro(self.h, level, rho.ctypes.data_as(dp), u.ctypes.data_as(dp)),
              "oal_download_macro")
        return rho, u

    # -- surface-stress forces (src/forces/surface.jl) ------------------------------------------------
    def set_surface(self, centers, normals, areas, level: int = 0) -> None:
        """Upload a triangle mesh (the reference's gpu_mesh): centers and normals [3, n_tri], areas
        [n_tri], Float32; None removes it."""
        if centers is None:
            check(self.lib.oal_set_surface(self.h, level, 0, None, None, None), "oal_set_surface")
            self._n_tri = getattr(self, "_n_tri", {})
            self._n_tri.pop(level, None)
            return
        c = np.ascontiguousarray(centers, dtype=np.float32)
        nrm = np.ascontiguousarray(normals, dtype=np.float32)
        a = np.ascontiguousarray(areas, dtype=np.float32)
        n = a.shape[0]
        if a.ndim != 1 or c.shape != (3, n) or nrm.shape != (3, n):
            raise ValueError("centers and normals must be [3, n_tri], areas [n_tri]")
        check(self.lib.oal_set_surface(self.h, level, n, c.ctypes.data, nrm.ctypes.data, a.ctypes.data),
              "oal_set_surface")
        self._n_tri = getattr(self, "_n_tri", {})
        self._n_tri[level] = n

    def surface_forces(self, level: int = 0, *, dx: float = 1.0, mesh_offset=(0.0, 0.0, 0.0),
                       pressure_scale: float = 1.0, stress_scale: float | None = None, tau: float | None = None,
                       c_wale: float = 0.2, nu_sgs_bg: float = 0.0, cs2: float = 1.0 / 3.0,
                       search_radius: int = 8, recover_normal: bool = True, symmetric: bool = False,
                       moment_center=(0.0, 0.0, 0.0), rho_ref: float = 1.0, u_ref: float = 0.0,
                       area_ref: float = 1.0, chord_ref: float = 1.0) -> dict:
        """map_surface_stresses! + integrate_surface_forces! (src/forces/surface.jl:1333-1548) on the
        level's current state: pressure and wall-law shear per triangle, integrated in FP64.  tau
        defaults to the level's first domain's; stress_scale to pressure_scale (as the reference sets
        it).  Returns the force_data fields: F_pressure, F_viscous, F, M (3-vectors), Cd, Cl, Cs, Cm
        and n_mapped."""
        p = _lib.OalSurfaceParams()
        p.dx = dx
        p.mesh_offset[:] = [float(v) for v in mesh_offset]
        p.pressure_scale = pressure_scale
        p.stress_scale = pressure_scale if stress_scale is None else stress_scale
        p.tau = self.levels[level].domains[0].tau if tau is None else tau
        p.c_wale, p.nu_sgs_bg, p.cs2 = c_wale, nu_sgs_bg, cs2
        p.moment_center[:] = [float(v) for v in moment_center]
        p.rho_ref, p.u_ref, p.area_ref, p.chord_ref = rho_ref, u_ref, area_ref, chord_ref
        p.search_radius, p.recover_normal, p.symmetric = int(search_radius), int(bool(recover_normal)), int(bool(symmetric))
        out = _lib.OalSurfaceForces()
        check(self.lib.oal_surface_forces(self.h, level, C.byref(p), C.byref(out), None), "oal_surface_forces")
        v3 = lambda a: np.array(list(a), dtype=np.float64)
        return {"F_pressure": v3(out.F_pressure), "F_viscous": v3(out.F_viscous), "F": v3(out.F), "M": v3(out.M),
                "Cd": out.Cd, "Cl": out.Cl, "Cs": out.Cs, "Cm": v3(out.Cm), "n_mapped": int(out.n_mapped)}

    def download_surface(self, level: int = 0) -> dict:
        """Per-triangle maps of the last surface_forces: pressure [n], shear [3, n], normal_sign [n],
        wall_dist [n] (Float32; pressure NaN where no fluid probe was found)."""
        n = getattr(self, "_n_tri", {}).get(level)
        if n is None:
            raise RuntimeError("no surface mesh on this level")
        p = np.empty(n, np.float32)
        s = np.empty((3, n), np.float32)
        sg = np.empty(n, np.float32)
        wd = np.empty(n, np.float32)
        check(self.lib.oal_download_surface(self.h, level, p.ctypes.data, s.ctypes.data, sg.ctypes.data,
                                            wd.ctypes.data), "oal_download_surface")
        return {"pressure": p, "shear": s, "normal_sign": sg, "wall_dist": wd}

    def guard(self, rho_min: float = 0.2, u_max: float = 0.5, level: int = 0) -> int:
        """Stability guard (src/octree_q19/src/io_run.jl:31-65); raises OalError when a cell is bad."""
        bad = C.c_int64()
        check(self.lib.oal_guard(self.h, level, rho_min, u_max, C.byref(bad)), "oal_guard")
        return bad.value


def perform_timestep_v2(plan: Plan, level: int = 0, timestep: int = 0, collect_force: bool = False) -> None:
    """One step of one level: the reference's perform_timestep_v2! (src/collision/timestep.jl:946-2000)
    with the operator of get_collision_kernel(:cumulant); asynchronous like the reference's launch."""
    plan.level_step(level, timestep, collect_force)


def execute_timestep_batch(plan: Plan, t_start: int, batch: int, collect_force_last: bool = False) -> int:
    """execute_timestep_batch! (src/solver/solver_control.jl:796-824): `batch` coarse steps from t_start,
    then a synchronisation; returns the next time step."""
    plan.step(batch, t_start, collect_force_last)
    plan.sync()
    return t_start + batch


def uniform_plan(grid: BlockGrid, tau: float, storage: str = "f32", stream: str = "ab", bb_mode: str = "halfway",
                 params: StepParams | None = None, u_inlet: float = 0.0, body_force=(0.0, 0.0, 0.0),
                 ymin_wall_bc: int = 0, zwall_bc: int = 0, rounding: str = "stochastic") -> Plan:
    """A single-level plan over one uniform domain covering the grid's extent."""
    dom = Domain(extent=grid.n, tau=tau, periodic=grid.periodic, u_inlet=u_inlet, body_force=tuple(body_force),
                 ymin_wall_bc=ymin_wall_bc, zwall_bc=zwall_bc)
    return Plan(Level(grid, [dom]), storage=storage, stream=stream, bb_mode=bb_mode, params=params, rounding=rounding)


def acoustic_taus(tau_finest: float, n_levels: int) -> list[float]:
    """Relaxation times of the levels (coarsest first) under acoustic scaling, dx and dt halving per level:
    tau_l - 1/2 = (tau_finest - 1/2) / 2^(L-1-l) (src/octree_q19/src/config.jl:328-335)."""
    return [0.5 + (tau_finest - 0.5) / 2.0 ** (n_levels - 1 - l) for l in range(n_levels)]


TAU_MIN = 0.501   # advanced.numerics.tau_min (src/io/config_loader.jl:83)


def level_taus(tau_finest: float, n_levels: int, rule: str = "legacy", tau_min: float = TAU_MIN) -> list[float]:
    """Relaxation times of the levels (coarsest first) by the GPU solver's rules
    (src/solver/physics_scaling.jl:654-682, advanced.refinement.level_viscosity):
      legacy (the default):  tau_l = 1/2 + (tau_f - 1/2) 2^(L-1-l);
      physical:              tau_l = max(1/2 + (tau_f - 1/2) / 2^(L-1-l), tau_min) below the finest level;
    the finest level keeps tau_f; each value is rounded to Float32 as the reference stores it."""
    if rule not in ("legacy", "physical"):
        raise ValueError("rule must be 'legacy' or 'physical'")
    out = []
    for l in range(n_levels):
        k = n_levels - 1 - l
        if k == 0:
            t = tau_finest
        elif rule == "legacy":
            t = 0.5 + (tau_finest - 0.5) * 2.0 ** k
        else:
            t = max(0.5 + (tau_finest - 0.5) / 2.0 ** k, float(np.float32(tau_min)))
        out.append(float(np.float32(t)))
    return out


def octree_plan(grids: list[BlockGrid], tau_finest: float, storage: str = "f32", params: StepParams | None = None,
                u_inlet: float = 0.0, body_force=(0.0, 0.0, 0.0), rounding: str = "stochastic",
                bb_mode: str = "halfway", taus: list[float] | None = None,
                amr_sgs_scales: list[float] | None = None) -> Plan:
    """A refined plan over the levels of a block octree (coarsest first, e.g. geometry.sphere_octree):
    one domain per level, by default with the acoustic-scaling relaxation times (taus: explicit per-level
    values instead, e.g. level_taus(..., "legacy")); the inlet velocity is the same in every level's
    lattice units.  amr_sgs_scales: each level's factor on the WALE and background eddy viscosity (by
    default 1 on every level, as the reference now sets it, src/collision/timestep.jl:1128; its earlier
    rule was 2^-(L-1-l))."""
    if taus is None:
        taus = acoustic_taus(tau_finest, len(grids))
    elif len(taus) != len(grids):
        raise ValueError("one tau per level")
    scales = [1.0] * len(grids) if amr_sgs_scales is None else list(amr_sgs_scales)
    if len(scales) != len(grids):
        raise ValueError("one amr_sgs_scale per level")
    levels = [Level(g, [Domain(extent=g.n, tau=t, periodic=g.periodic, u_inlet=u_inlet, body_force=tuple(body_force),
                               amr_sgs_scale=a)])
              for g, t, a in zip(grids, taus, scales)]
    return Plan(levels, storage=storage, stream="ab", bb_mode=bb_mode, params=params, rounding=rounding)


def batched_plan(grids: list[BlockGrid], taus, u_inlets, storage: str = "f16s", params: StepParams | None = None,
                 rounding: str = "stochastic", bb_mode: str = "halfway") -> Plan:
    """Independent boxes (inlet / outlet / free-slip channels) advanced together: one level whose block
    list holds every box (geometry.batch_grids), one domain per box with its own relaxation time and inlet
    velocity.  Every launch of a step covers all boxes; force and mass are per box (per domain)."""
    from .geometry import batch_grids
    big = batch_grids(grids)
    doms = [Domain(extent=g.n, tau=float(t), u_inlet=float(u)) for g, t, u in zip(grids, taus, u_inlets)]
    return Plan(Level(big, doms, block_domain=big.block_domain), storage=storage, stream="ab", bb_mode=bb_mode,
                params=params, rounding=rounding)
