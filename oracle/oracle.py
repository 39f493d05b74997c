"""ctypes wrapper of the CPU oracle (oracle/liboal_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker; the product (openaerolab_alpha_01_amd) never imports it.

Parity status of each oracle function and what pins it (tests/test_oracle_*.py).  The pins of
test_oracle_pins.py each come with a planted mistake (`planted`, a test hook of the C++ oracle) that the
same check is shown to catch.

  lattice tables        pinned: isotropy identities (sum w = 1, sum w c c = I/3, 4th-order isotropy),
                        opp / mirror involutions, index formula of lattice.jl:16-19
  f16 / bf16 rounding   pinned: exhaustive over all 2^16 codes (f16 -> f32 -> f16 round trip) and
                        against numpy's float16 cast (round to nearest even) on random floats
  equilibria            pinned: exact moments rho, rho u, rho (u u + I/3) (2nd order); 3rd/4th-order
                        Hermite moments of the O(u^4) form
  inlet noise hash      pinned (test_oracle_pins.py): gpu_hash is MurmurHash3_x86_32's finaliser, checked
                        against 13 published MurmurHash3_x86_32 test vectors (empty-key vectors are the
                        finaliser of the seed); gradient_noise's Int32 combination wraps modulo 2^32 on
                        overflowing coordinates; bijectivity and range [-1, 1) (test_oracle_lattice.py)
  chimera transforms    pinned: forward equals brute-force central moments sum f (c-u)^a (FP64);
                        backward(forward(f)) = f
  5th/6th-order rebuild pinned (test_oracle_pins.py): the post-collision 5th- and 6th-order cumulants,
                        computed from brute-force moments by the set-partition formula, vanish with
                        omega_3, omega_4 != 1 and 3rd-order input
  collision (one cell)  pinned: mass and momentum conservation (no force); rest state fixed point;
                        omega1 = 1 with omega_bulk = omega_3 = omega_4 = 1 gives the cumulant
                        equilibrium (2nd-order cumulants I/3, higher cumulants zero) checked in moment
                        space by brute force; relaxation of the deviatoric 2nd-order cumulants by
                        (1 - omega1) checked by brute-force cumulants
  step (periodic)       pinned: shear-wave decay nu = (tau - 1/2)/3 (1 %), Taylor-Green KE decay
                        2nd-order convergence, mass/momentum conservation, rest state
  halfway bounce-back   pinned: Poiseuille channel, momentum exchange = body force x fluid volume,
  + momentum exchange   velocity profile vs the parabola
  domain-face rules     pinned: uniform inflow through inlet/outlet/free-slip stays uniform; mirror
                        rules by symmetry of a mirrored run
  WALE eddy viscosity   pinned: the model's tensor definition evaluated independently (einsum), zero in
                        pure shear, degree-1 homogeneity, frame invariance (test_oracle_wale.py)
  WALE gradient gather  pinned (test_oracle_pins.py): exact gradient of a linear field with fluid
                        neighbours; one-sided half differences next to solid cells, the extent and an
                        absent block, derived by hand; periodic wrap
  refined runs          pinned (octree.py, test_oracle_octree.py): cubic weights reproduce cubics,
                        stencils are partitions of unity and reproduce linear fields, uniform flow
                        across 2 and 3 levels, restriction of the prolongation on a freestream, shear
                        wave across a refined slab at the physical viscosity with composite mass drift
                        under 1e-4; the exact rounding of the transfers is parity unpinned (no
                        reference vector)
  interpolated          pinned (test_oracle_bfl.py): q = 1/2 reproduces halfway bounce-back bit for bit,
  bounce-back (BFL)     with the gather-time force equal to the halfway force of the step before;
                        Poiseuille walls off the half-link (q = 1/4, 3/4) follow the parabola of the
                        q-placed walls; rest state kept under every fallback; fallbacks 0, 1 and 2 away
                        from rest on a single cut link, the arriving value read back from the link's
                        momentum exchange and derived by hand (test_oracle_pins.py)
  sponge blend          pinned (test_oracle_sponge.py): mass and momentum after the collision equal the
                        blended state's (rho (1 - sp) + sp, and the blended velocity), in both blend modes
                        and both arithmetic forms; sp = 0 is no sponge bit for bit
  face-rule variants    pinned (test_oracle_faces.py): exact specular z walls equal one step of the box
                        doubled in z by its mirror image (bit for bit); the transverse freestream's face
                        cells hold the state's moments plus b times the change of the crossing
                        populations (b = 0, 0.3, 1) and keep the freestream equilibrium
  surface stress        pinned (test_oracle_surface.py): uniform pressure gives zero net force and
                        moment on a closed mesh; uniform flow's wall-law shear integrates to the sphere's
                        closed form with and without the SGS estimate; a linear density gives the
                        divergence-theorem force and the moment arm; hand-derived single triangles for
                        the quality gate, the opposite-side recovery and the shell search
  golden fixtures       tests/golden/ (written from this oracle; regression pins, test_golden.py)
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboal_oracle.so"

STORAGE = {"f32": 0, "f64": 1, "f16s": 2, "bf16s": 3}
LITERAL, FAST = 0, 1
STOCHASTIC, NEAREST = 0, 1
F16_SCALE = np.float32(16384.0)   # FP16 storage holds (f - w) * 2^14
STORE_DTYPE = {"f32": np.float32, "f64": np.float64, "f16s": np.uint16, "bf16s": np.uint16}
WORK_DTYPE = {"f32": np.float32, "f64": np.float64, "f16s": np.float32, "bf16s": np.float32}


class OrParams(C.Structure):
    _fields_ = [
        ("tau", C.c_double), ("omega_bulk", C.c_double), ("omega3", C.c_double), ("omega4", C.c_double),
        ("lambda_", C.c_double), ("c_wale", C.c_double), ("nu_sgs_bg", C.c_double), ("amr_sgs_scale", C.c_double),
        ("u_inlet", C.c_double), ("inlet_turb", C.c_double), ("body_force", C.c_double * 3),
        ("adaptive_omega4", C.c_int32), ("limiter", C.c_int32), ("compressible", C.c_int32),
        ("ymin_wall_bc", C.c_int32), ("zwall_bc", C.c_int32), ("bb_mode", C.c_int32), ("time_step_seed", C.c_int32),
        ("arith", C.c_int32), ("rounding", C.c_int32), ("bfl_fallback", C.c_int32), ("q_min", C.c_double),
        ("sponge_blend", C.c_int32), ("transverse_freestream", C.c_int32), ("tfs_blend", C.c_double),
    ]


class OrGrid(C.Structure):
    _fields_ = [
        ("bs", C.c_int32), ("nb", C.c_int32), ("bcoord", C.POINTER(C.c_int32)), ("nbr", C.POINTER(C.c_int32)),
        ("n", C.c_int32 * 3), ("periodic", C.c_int32 * 3), ("solid", C.POINTER(C.c_uint8)),
        ("active", C.POINTER(C.c_uint8)), ("qmap", C.POINTER(C.c_uint16)), ("sponge", C.POINTER(C.c_float)),
    ]


class OrSurf(C.Structure):
    _fields_ = [
        ("dx", C.c_double), ("mesh_offset", C.c_double * 3), ("pressure_scale", C.c_double),
        ("stress_scale", C.c_double), ("tau", C.c_double), ("c_wale", C.c_double), ("nu_sgs_bg", C.c_double),
        ("cs2", C.c_double), ("moment_center", C.c_double * 3), ("search_radius", C.c_int32),
        ("recover_normal", C.c_int32),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        subprocess.run(["make", "-C", str(HERE), "liboal_oracle.so"], check=True, stdout=subprocess.DEVNULL)
    lib = C.CDLL(str(LIB))
    vp = C.c_void_p
    lib.oracle_step.argtypes = [C.c_int, C.POINTER(OrGrid), C.POINTER(OrParams), vp, vp, vp, vp, vp, C.POINTER(C.c_double)]
    lib.oracle_step.restype = C.c_int
    lib.oracle_init_eq.argtypes = [C.c_int, C.POINTER(OrGrid), vp, vp, vp]
    lib.oracle_moments.argtypes = [C.c_int, C.POINTER(OrGrid), vp, C.POINTER(C.c_double), vp, vp]
    lib.oracle_lattice.argtypes = [vp] * 8
    lib.oracle_gradient_noise.argtypes = [C.c_int32] * 4
    lib.oracle_gradient_noise.restype = C.c_float
    lib.oracle_gpu_hash.argtypes = [C.c_int32]
    lib.oracle_gpu_hash.restype = C.c_uint32
    lib.oracle_feq_f64.argtypes = [C.c_double] * 4 + [C.c_int, C.c_int]
    lib.oracle_feq_f64.restype = C.c_double
    lib.oracle_feq_f32.argtypes = [C.c_float] * 4 + [C.c_int, C.c_int]
    lib.oracle_feq_f32.restype = C.c_float
    lib.oracle_chimera_f64.argtypes = [vp, C.c_double, C.c_int]
    lib.oracle_collide_f64.argtypes = [vp, C.POINTER(OrParams), C.c_double, vp, vp]
    lib.oracle_collide_f32.argtypes = [vp, C.POINTER(OrParams), C.c_float, vp, vp]
    lib.oracle_collide_fast_f64.argtypes = [vp, C.POINTER(OrParams), C.c_double, vp, vp]
    lib.oracle_collide_fast_f32.argtypes = [vp, C.POINTER(OrParams), C.c_float, vp, vp]
    lib.oracle_wale_nu_f64.argtypes = [vp, C.c_double]
    lib.oracle_wale_nu_f64.restype = C.c_double
    for nm in ("oracle_f32_to_f16", "oracle_f16_to_f32", "oracle_f32_to_bf16", "oracle_bf16_to_f32"):
        getattr(lib, nm).argtypes = [vp, vp, C.c_int64]
    for nm in ("oracle_f32_to_f16_sr", "oracle_f32_to_bf16_sr"):
        getattr(lib, nm).argtypes = [vp, vp, vp, C.c_int64]
    lib.oracle_round_bits.argtypes = [vp, vp, C.c_int32, vp, C.c_int64]
    lib.oracle_set_plant.argtypes = [C.c_int]
    lib.oracle_vel_gradients.argtypes = [C.POINTER(OrGrid), vp, vp]
    lib.oracle_surface.argtypes = [C.POINTER(OrGrid), vp, C.c_int, C.POINTER(OrSurf), C.c_int64, vp, vp, vp, vp, vp,
                                   vp, vp, vp]
    lib.oracle_surface.restype = C.c_int
    _lib = lib
    return lib


def f32(v: float) -> float:
    """A parameter as the library receives it (Float32 kernel arguments of the reference)."""
    return float(np.float32(v))


def make_params(tau, *, omega_bulk=1.0, omega_3=1.0, omega_4=1.0, lambda_=1.0 / 6.0, adaptive_omega4=1, limiter=1,
                compressible=0, c_wale=0.0, nu_sgs_background=0.0, amr_sgs_scale=1.0, u_inlet=0.0,
                inlet_turbulence=0.0, body_force=(0.0, 0.0, 0.0), ymin_wall_bc=0, zwall_bc=0, bb_mode=1,
                time_step=0, arith=LITERAL, rounding=STOCHASTIC, bfl_fallback=0, q_min=0.001,
                sponge_blend=0, transverse_freestream=0, transverse_freestream_blend=1.0) -> OrParams:
    """Step parameters; arith LITERAL restates the reference arithmetic operation for operation, FAST the
    library's fused arithmetic (same mathematics, fused multiply-adds, rho and j from the transform's
    partial sums).  rounding: how the 16-bit forms store a step's populations (STOCHASTIC, the
    library's default, or NEAREST).  bfl_fallback and q_min apply to a grid with a q-map (interpolated
    bounce-back): the rule of a q < 1/2 link with no fluid cell behind (0 halfway, 1 wall equilibrium,
    2 own opposite population) and the smallest q that is interpolated."""
    p = OrParams()
    p.tau, p.omega_bulk, p.omega3, p.omega4 = f32(tau), f32(omega_bulk), f32(omega_3), f32(omega_4)
    p.lambda_, p.c_wale, p.nu_sgs_bg, p.amr_sgs_scale = f32(lambda_), f32(c_wale), f32(nu_sgs_background), f32(amr_sgs_scale)
    p.u_inlet, p.inlet_turb = f32(u_inlet), f32(inlet_turbulence)
    p.body_force[:] = [f32(v) for v in body_force]
    p.adaptive_omega4, p.limiter, p.compressible = adaptive_omega4, limiter, compressible
    p.ymin_wall_bc, p.zwall_bc, p.bb_mode, p.time_step_seed = ymin_wall_bc, zwall_bc, bb_mode, time_step
    p.arith = arith
    p.rounding = rounding
    p.bfl_fallback = bfl_fallback
    p.q_min = f32(q_min)
    p.sponge_blend = sponge_blend
    p.transverse_freestream, p.tfs_blend = transverse_freestream, f32(transverse_freestream_blend)
    return p


class Grid:
    """The oracle's view of a block grid (any object with n, periodic, coords, neighbor, solid, bs).
    qmap: [27, n_cells] float16 wall distances of the links (interpolated bounce-back), or None.
    sponge: [n_cells] sponge strength per cell (float32), or None."""

    def __init__(self, g, qmap=None, sponge=None):
        self.src = g
        self.bs = int(g.bs)
        self.nb = int(g.coords.shape[0])
        self.ncell = self.nb * self.bs ** 3
        self.bcoord = np.ascontiguousarray(np.asarray(g.coords, dtype=np.int32) - 1)
        # dir-major 0-based table: nbr[d*nb + b]
        self.nbr = np.ascontiguousarray(np.asarray(g.neighbor, dtype=np.int32).T - 1)
        self.solid = None if g.solid is None else np.ascontiguousarray(g.solid, dtype=np.uint8)
        s = OrGrid()
        s.bs, s.nb = self.bs, self.nb
        s.bcoord = self.bcoord.ctypes.data_as(C.POINTER(C.c_int32))
        s.nbr = self.nbr.ctypes.data_as(C.POINTER(C.c_int32))
        s.n[:] = [int(v) for v in g.n]
        s.periodic[:] = [1 if p else 0 for p in g.periodic]
        s.solid = None if self.solid is None else self.solid.ctypes.data_as(C.POINTER(C.c_uint8))
        act = getattr(g, "active", None)
        self.active = None if act is None else np.ascontiguousarray(act, dtype=np.uint8)
        s.active = None if self.active is None else self.active.ctypes.data_as(C.POINTER(C.c_uint8))
        self.qmap = None
        if qmap is not None:
            q = np.asarray(qmap)
            if q.dtype != np.uint16:
                q = q.astype(np.float16).view(np.uint16)
            if q.shape != (27, self.ncell):
                raise ValueError(f"qmap must be [27, {self.ncell}]")
            self.qmap = np.ascontiguousarray(q)
        s.qmap = None if self.qmap is None else self.qmap.ctypes.data_as(C.POINTER(C.c_uint16))
        self.sponge = None if sponge is None else np.ascontiguousarray(sponge, dtype=np.float32)
        if self.sponge is not None and self.sponge.shape != (self.ncell,):
            raise ValueError(f"sponge must be [{self.ncell}]")
        s.sponge = None if self.sponge is None else self.sponge.ctypes.data_as(C.POINTER(C.c_float))
        self.c = s


def init_eq(storage: str, grid: Grid, rho: np.ndarray, u: np.ndarray) -> np.ndarray:
    lib = load()
    wt = WORK_DTYPE[storage]
    r = np.ascontiguousarray(rho, dtype=wt)
    v = np.ascontiguousarray(u, dtype=wt)
    f = np.zeros((27, grid.ncell), dtype=STORE_DTYPE[storage])
    lib.oracle_init_eq(STORAGE[storage], C.byref(grid.c), r.ctypes.data, v.ctypes.data, f.ctypes.data)
    return f


def step(storage: str, grid: Grid, params: OrParams, fin: np.ndarray, vel_in: np.ndarray | None = None,
         want_macro: bool = False, want_force: bool = False, out: np.ndarray | None = None):
    """One step; returns (fout, force or None, (rho, vel) or None).  fout starts as a copy of fin so
    that cells the step does not write (solid, beyond the extent) keep their values; a caller-provided
    `out` is used as is (the timing loop of the CPU baseline alternates two buffers)."""
    lib = load()
    fout = fin.copy() if out is None else out
    wt = WORK_DTYPE[storage]
    rho = np.zeros(grid.ncell, dtype=wt) if want_macro else None
    vel = np.zeros((3, grid.ncell), dtype=wt) if want_macro else None
    vin = None if vel_in is None else np.ascontiguousarray(vel_in, dtype=wt)
    F = (C.c_double * 3)()
    rc = lib.oracle_step(STORAGE[storage], C.byref(grid.c), C.byref(params), fin.ctypes.data, fout.ctypes.data,
                         None if vin is None else vin.ctypes.data, None if vel is None else vel.ctypes.data,
                         None if rho is None else rho.ctypes.data, F if want_force else None)
    assert rc == 0
    return fout, (np.array(F[:]) if want_force else None), ((rho, vel) if want_macro else None)


def moments(storage: str, grid: Grid, f: np.ndarray):
    """(mass, jx, jy, jz) in FP64 and per-cell rho [ncell], u [3, ncell] (FP64)."""
    lib = load()
    out = (C.c_double * 4)()
    rho = np.zeros(grid.ncell)
    u = np.zeros((3, grid.ncell))
    lib.oracle_moments(STORAGE[storage], C.byref(grid.c), f.ctypes.data, out, rho.ctypes.data, u.ctypes.data)
    return np.array(out[:]), rho, u


def vel_gradients(grid: Grid, vel: np.ndarray) -> np.ndarray:
    """WALE's velocity gradients [3, 3, ncell] (g[a, j] = d u_a / d x_j) of a velocity field [3, ncell]."""
    v = np.ascontiguousarray(vel, dtype=np.float64)
    out = np.zeros((9, grid.ncell))
    load().oracle_vel_gradients(C.byref(grid.c), v.ctypes.data, out.ctypes.data)
    return out.reshape(3, 3, grid.ncell)


def surface(grid: Grid, f: np.ndarray, centers, normals, areas, *, w32: bool = True, dx: float = 1.0,
            mesh_offset=(0.0, 0.0, 0.0), pressure_scale: float = 1.0, stress_scale: float | None = None,
            tau: float = 0.6, c_wale: float = 0.2, nu_sgs_bg: float = 0.0, cs2: float = 1.0 / 3.0,
            search_radius: int = 8, recover_normal: bool = True, moment_center=(0.0, 0.0, 0.0)) -> dict:
    """Surface-stress map and forces (src/forces/surface.jl) of a state given as unshifted populations
    f [27, ncell] (FP64 values; w32: the state's working precision is FP32).  Returns the per-triangle
    maps (pressure, shear [3, n], normal_sign, wall_dist; Float32) and the FP64 sums F_pressure,
    F_viscous, M (3-vectors) and n_mapped."""
    lib = load()
    ff = np.ascontiguousarray(f, dtype=np.float64)
    c = np.ascontiguousarray(centers, dtype=np.float32)
    nrm = np.ascontiguousarray(normals, dtype=np.float32)
    a = np.ascontiguousarray(areas, dtype=np.float32)
    n = a.shape[0]
    s = OrSurf()
    s.dx = dx
    s.mesh_offset[:] = [float(v) for v in mesh_offset]
    s.pressure_scale = pressure_scale
    s.stress_scale = pressure_scale if stress_scale is None else stress_scale
    s.tau, s.c_wale, s.nu_sgs_bg, s.cs2 = tau, c_wale, nu_sgs_bg, cs2
    s.moment_center[:] = [float(v) for v in moment_center]
    s.search_radius, s.recover_normal = int(search_radius), int(bool(recover_normal))
    p = np.empty(n, np.float32)
    sh = np.empty((3, n), np.float32)
    sg = np.empty(n, np.float32)
    wd = np.empty(n, np.float32)
    sums = np.zeros(10)
    rc = lib.oracle_surface(C.byref(grid.c), ff.ctypes.data, int(bool(w32)), C.byref(s), n, c.ctypes.data,
                            nrm.ctypes.data, a.ctypes.data, p.ctypes.data, sh.ctypes.data, sg.ctypes.data,
                            wd.ctypes.data, sums.ctypes.data)
    assert rc == 0
    return {"pressure": p, "shear": sh, "normal_sign": sg, "wall_dist": wd, "F_pressure": sums[0:3].copy(),
            "F_viscous": sums[3:6].copy(), "M": sums[6:9].copy(), "n_mapped": int(sums[9])}


class planted:
    """Context manager that switches on one of the oracle's planted mistakes (test hook; see
    oal_oracle.cpp, g_plant), so that a pin test can show its check fails on it."""

    def __init__(self, which: int):
        self.which = which

    def __enter__(self):
        load().oracle_set_plant(self.which)
        return self

    def __exit__(self, *exc):
        load().oracle_set_plant(0)
        return False


def to_unshifted(storage: str, f: np.ndarray) -> np.ndarray:
    """Stored populations -> unshifted float64 [27, ncell]."""
    if storage in ("f32", "f64"):
        return f.astype(np.float64)
    lib = load()
    flat = np.ascontiguousarray(f).ravel()
    out = np.empty(flat.size, dtype=np.float32)
    fn = lib.oracle_f16_to_f32 if storage == "f16s" else lib.oracle_bf16_to_f32
    fn(flat.ctypes.data, out.ctypes.data, flat.size)
    if storage == "f16s":
        out *= np.float32(1.0) / F16_SCALE
    w = lattice()[6]
    return (out.reshape(f.shape) + w.astype(np.float32)[:, None]).astype(np.float64)


def from_unshifted(storage: str, f: np.ndarray) -> np.ndarray:
    """Unshifted populations [27, ncell] -> the storage form (rounded to nearest even)."""
    if storage == "f32":
        return np.ascontiguousarray(f, dtype=np.float32)
    if storage == "f64":
        return np.ascontiguousarray(f, dtype=np.float64)
    lib = load()
    w = lattice()[6]
    s = np.ascontiguousarray((f.astype(np.float32) - w.astype(np.float32)[:, None]).astype(np.float32))
    if storage == "f16s":
        s *= F16_SCALE
    out = np.empty(s.shape, dtype=np.uint16)
    fn = lib.oracle_f32_to_f16 if storage == "f16s" else lib.oracle_f32_to_bf16
    fn(s.ctypes.data, out.ctypes.data, s.size)
    return out


def f32_to_f16_sr(x: np.ndarray, r: np.ndarray) -> np.ndarray:
    """FP32 -> FP16 codes, stochastic rounding with the 32-bit random words r."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    r = np.ascontiguousarray(r, dtype=np.uint32)
    out = np.empty(x.shape, dtype=np.uint16)
    load().oracle_f32_to_f16_sr(x.ctypes.data, r.ctypes.data, out.ctypes.data, x.size)
    return out


def f32_to_bf16_sr(x: np.ndarray, r: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 codes, stochastic rounding with the 32-bit random words r."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    r = np.ascontiguousarray(r, dtype=np.uint32)
    out = np.empty(x.shape, dtype=np.uint16)
    load().oracle_f32_to_bf16_sr(x.ctypes.data, r.ctypes.data, out.ctypes.data, x.size)
    return out


def round_bits(cell: np.ndarray, k: np.ndarray, time_step: int) -> np.ndarray:
    """The random words of the stores of population k of cell `cell` at a time step."""
    cell = np.ascontiguousarray(cell, dtype=np.int64)
    k = np.ascontiguousarray(k, dtype=np.int32)
    out = np.empty(cell.shape, dtype=np.uint32)
    load().oracle_round_bits(cell.ctypes.data, k.ctypes.data, int(time_step), out.ctypes.data, cell.size)
    return out


def lattice():
    """(cx, cy, cz, opp, mir_y, mir_z, w32, w64) of the oracle."""
    lib = load()
    a = [np.zeros(27, dtype=np.int32) for _ in range(6)]
    w32 = np.zeros(27, dtype=np.float32)
    w64 = np.zeros(27, dtype=np.float64)
    lib.oracle_lattice(*[x.ctypes.data for x in a], w32.ctypes.data, w64.ctypes.data)
    return (*a, w32, w64)


def collide_cell_f64(f: np.ndarray, params: OrParams, nu_eddy: float = 0.0):
    """One fluid-cell collision in FP64, in the arithmetic form params.arith selects."""
    lib = load()
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(27)
    ru = np.zeros(4)
    fn = lib.oracle_collide_fast_f64 if params.arith == FAST else lib.oracle_collide_f64
    fn(f.ctypes.data, C.byref(params), nu_eddy, out.ctypes.data, ru.ctypes.data)
    return out, ru


def collide_cell_f32(f: np.ndarray, params: OrParams, nu_eddy: float = 0.0):
    lib = load()
    f = np.ascontiguousarray(f, dtype=np.float32)
    out = np.zeros(27, dtype=np.float32)
    ru = np.zeros(4, dtype=np.float32)
    fn = lib.oracle_collide_fast_f32 if params.arith == FAST else lib.oracle_collide_f32
    fn(f.ctypes.data, C.byref(params), nu_eddy, out.ctypes.data, ru.ctypes.data)
    return out, ru


def chimera_f64(abc, u: float, direction: int) -> np.ndarray:
    lib = load()
    a = np.ascontiguousarray(abc, dtype=np.float64).copy()
    lib.oracle_chimera_f64(a.ctypes.data, u, direction)
    return a
